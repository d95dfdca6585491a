"""Batched serving: every result is the same alone and in the mix.

The batched step kernel runs each solve's strips with the single-solve
arithmetic, so the probe traces and final fields of a solve served in a mix of
different grid sizes, admitted at different times and finishing at different
steps, equal (bit for bit) those of the same request run alone.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_requests(n=6, seed=7):
    from strata_fdtd_amd.serving import sweep_requests
    reqs = sweep_requests(n=n, rate=1e6, seed=seed, steps=(40, 130), dims=(64, 80))
    return reqs


@pytest.mark.parametrize("chunk,use_graph,rows", [(16, True, 4), (7, False, 4), (16, True, 2)])
def test_served_results_equal_solo_runs(chunk, use_graph, rows):
    from strata_fdtd_amd.serving import SweepServer, build_solver
    reqs = _small_requests()
    print(f"[serving] {len(reqs)} requests, chunk={chunk}, graph={use_graph}, rows={rows}", flush=True)
    srv = SweepServer(chunk=chunk, use_graph=use_graph, rows=rows)
    stats = srv.serve(reqs, realtime=False)
    srv.close()
    assert len(stats.results) == len(reqs)
    assert stats.max_active > 1
    got = {r.rid: r.traces for r in stats.results}
    for r in reqs:
        s = build_solver(r, tiling=(4, 2))
        s.step_n(r.steps, graph_chunk=0)
        alone = s.get_probe_data()
        for name, tr in alone.items():
            assert len(got[r.rid][name]) == r.steps
            assert np.array_equal(got[r.rid][name], tr), (r.rid, name)


def test_admission_under_tight_memory_budget():
    """A budget that fits about two solves forces queueing; results are unchanged."""
    from strata_fdtd_amd.serving import SweepServer, build_solver
    reqs = _small_requests(n=5, seed=11)
    budget = 2 * max(r.footprint for r in reqs)
    srv = SweepServer(chunk=10, hbm_budget=budget)
    stats = srv.serve(reqs, realtime=False)
    srv.close()
    assert len(stats.results) == len(reqs)
    assert stats.max_active <= 2
    for res in stats.results:
        r = next(x for x in reqs if x.rid == res.rid)
        s = build_solver(r)
        s.step_n(r.steps)
        for name, tr in s.get_probe_data().items():
            assert np.array_equal(res.traces[name], tr)


def test_batch_beyond_the_cache_equals_solo_runs():
    """Three 160^3 solves (12M cells together, beyond the Infinity-Cache
    working set: the batch is written with non-temporal stores) admitted at
    once, FIFO: every result equals the solo run's."""
    from strata_fdtd_amd.serving import SweepServer, build_solver, sweep_requests
    reqs = sweep_requests(n=3, rate=1e6, seed=3, steps=(150, 220), dims=(160,))
    srv = SweepServer(chunk=16, max_cells=None, policy="fifo")
    stats = srv.serve(reqs, realtime=False)
    srv.close()
    assert stats.max_active == 3
    for res in stats.results:
        r = next(x for x in reqs if x.rid == res.rid)
        s = build_solver(r)
        s.step_n(r.steps)
        for name, tr in s.get_probe_data().items():
            assert np.array_equal(res.traces[name], tr)


@pytest.mark.parametrize("rows", [2, 4])
def test_batch_fields_equal_solo_runs(rows):
    """Solves whose k extent leaves a partial last 64-lane segment of 16, 32
    or 48 cells (nz = 80, 96, 112, 144), every cell starting nonzero: after
    steps in chunks of 7 (the last chunk partial for every solve), all four
    final fields and the probe traces of every solve equal its solo run bit
    for bit, and the step counters the chunks moved stop at each limit."""
    import torch
    from strata_fdtd_amd import _native
    from strata_fdtd_amd.serving import build_solver, sweep_requests
    # nz = 96, 80, 144 and 112
    reqs = sweep_requests(n=4, rate=1e6, seed=5, steps=(30, 50), dims=(80, 96, 112, 144))
    print(f"[serving] batch fields, rows={rows}: " + ", ".join(f"{r.shape} x{r.steps}" for r in reqs), flush=True)

    def seeded(r, q):
        # every cell starts nonzero, so every cell's arithmetic is compared
        s = build_solver(r, tiling=(rows, 2, 1))
        g = np.random.default_rng(100 + q)
        for k in ("p", "vx", "vy", "vz"):
            getattr(s, k)[...] = g.standard_normal(r.shape).astype(np.float32)
        return s

    batch = [seeded(r, q) for q, r in enumerate(reqs)]
    for s, r in zip(batch, reqs):
        s._prepare(r.steps)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    h = _native._vp()
    _native.call("fdtd_batch_create", 0, stream.cuda_stream, rows, _native.C.byref(h))
    try:
        ctxs = (_native._vp * len(batch))(*[s._ctx for s in batch])
        limits = np.array([r.steps for r in reqs], dtype=np.int64)
        _native.call("fdtd_batch_set", h.value, ctxs, limits.ctypes.data, len(batch))
        for _ in range(0, max(r.steps for r in reqs), 7):
            _native.call("fdtd_batch_step_n", h.value, 7, 0)
        torch.cuda.synchronize()
    finally:
        _native.load().fdtd_batch_destroy(h.value)
    for s, r in zip(batch, reqs):
        s._advance_host(r.steps)
        solo = seeded(r, reqs.index(r))
        solo.step_n(r.steps)
        assert int(s._step_dev.item()) == r.steps
        for k in ("p", "vx", "vy", "vz"):
            a, b = getattr(s, k), getattr(solo, k)
            assert np.array_equal(a, b), (r.shape, k, int((a != b).sum()))
        for name, tr in solo.get_probe_data().items():
            assert np.array_equal(s.get_probe_data()[name], tr), (r.shape, name)
