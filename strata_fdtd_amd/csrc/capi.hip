// C ABI of the MI355X FDTD backend (declarations in include/fdtd_mi355x.h).
// Host-side only: validates arguments, keeps the launch arguments of one solve
// in a context, enqueues kernels on the context stream. No entry point that
// steps the solver allocates device memory or synchronises.

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fdtd_common.hpp"

namespace fdtd_amd {
int fused_grid(FusedArgs& a, int rows, int planes, int stack);
hipError_t launch_fused(const FusedArgs& a, int rows, int stack, bool nu, bool ade, int blocks,
                        hipStream_t st, hipEvent_t ev_before, hipEvent_t ev_after, const BoundarySet* bnd);
hipError_t launch_advance(int64_t* step, int n, hipStream_t st);
hipError_t launch_stream_ceiling(const FusedArgs& a, hipStream_t st);
hipError_t launch_stream_march(const FusedArgs& a, int rows, int planes, hipStream_t st);
int energy_partials();
hipError_t launch_energy(const FusedArgs& a, double* partial, double* out, hipStream_t st);
hipError_t launch_centre_velocity(const FusedArgs& a, float* ox, float* oy, float* oz, hipStream_t st);
hipError_t launch_batch(const FusedArgs* table, const int* offsets, const int* n_ptr, int cap_waves, int rows,
                        bool nts, int koff, hipStream_t st);
hipError_t launch_batch_advance(const FusedArgs* table, const int* n_ptr, int cap_solves, int n, hipStream_t st);
hipError_t phase_velocity(const float*, float*, float*, float*, int, int, int, float, const float*,
                          const float*, const float*, hipStream_t);
hipError_t phase_rigid(float*, float*, float*, const uint8_t*, int, int, int, hipStream_t);
hipError_t phase_pressure(float*, const float*, const float*, const float*, const uint8_t*, int, int,
                          int, float, const float*, const float*, const float*, hipStream_t);
hipError_t phase_pml_velocity(float*, float*, float*, const float*, const float*, const float*, int,
                              int, int, hipStream_t);
hipError_t phase_pml_pressure(float*, const float*, const float*, const float*, int, int, int,
                              hipStream_t);
hipError_t phase_divergence(float*, const float*, const float*, const float*, int, int, int, float,
                            const float*, const float*, const float*, hipStream_t);
hipError_t phase_build_flags(const uint8_t*, const uint8_t*, uint8_t*, int, int, int, hipStream_t);
hipError_t phase_ade_update(float* J, float* Jp, const float* src, const uint8_t* mat, int mat_id, int lorentz,
                            float a, float b, float d, int64_t n, hipStream_t st);
hipError_t phase_ade_vcorr(float* vx, float* vy, float* vz, const float* J, const uint8_t* mat, int mat_id,
                           float coeff, int nx, int ny, int nz, hipStream_t st);
hipError_t phase_ade_pcorr(float* p, const float* J, const uint8_t* mat, int mat_id, float coeff, int64_t n,
                           hipStream_t st);
struct VoxProgram {
    const int32_t* ops;
    const int32_t* start;
    const int32_t* ids;
    const double* params;
    int n_prog;
};
const char* vox_validate(const int32_t* ops, int n_ops, const int32_t* start, int n_prog, const int32_t* ids,
                         const double* params, int n_params);
hipError_t launch_voxelize(const VoxProgram& pg, const double* xc, const double* yc, const double* zc, int nx,
                           int ny, int nz, uint8_t* air, uint8_t* ids, hipStream_t st);
}  // namespace fdtd_amd

using namespace fdtd_amd;

namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail(-5, "%s: %s", what, hipGetErrorString(e));
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};
}  // namespace

struct fdtd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    FusedArgs a{};
    bool nonuniform = false;
    bool have_grid = false, have_fields = false, have_flags = false, have_pml = false;
    bool have_counters = false;
    int rows = 4, planes = 8, stack = 1;
    int32_t* mic_idx_dev = nullptr;   // owned: [mic_cap][4][8]
    double* mic_w_dev = nullptr;      // owned: [mic_cap][4][8]
    int32_t* mic_f64_dev = nullptr;   // owned: [mic_cap][4]
    int mic_cap = 0;
    // owned membrane tables: one allocation each, grown as needed
    struct DevBuf {
        void* p = nullptr;
        size_t bytes = 0;
    } m_cell, m_estart, m_esrc, m_efield, m_eweight, m_pstart, m_pidx, ade_cells, ade_poles, probes;
    int64_t* step_counter = nullptr;  // same as a.step_ptr, writable
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> sweep_ev;  // pairs around each sweep (fdtd_time_sweeps)
    double* energy_partial = nullptr;  // owned: [energy_partials()][2]
    BoundarySet bnd;                   // boundary planes (state owned by the caller)
};

extern "C" {

const char* fdtd_last_error(void) { return g_err.c_str(); }

const char* fdtd_version(void) { return "strata_fdtd_amd 0.1.0 (gfx950, HIP)"; }

int fdtd_device_count(int* out) {
    if (!out) return fail(-1, "fdtd_device_count: null out");
    return check(hipGetDeviceCount(out), "hipGetDeviceCount");
}

int fdtd_ctx_create(int device, void* stream, fdtd_ctx** out) {
    if (!out) return fail(-1, "fdtd_ctx_create: null out");
    int n = 0;
    if (int rc = check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return rc;
    if (device < 0 || device >= n) return fail(-1, "fdtd_ctx_create: device %d of %d", device, n);
    auto* c = new fdtd_ctx();
    c->device = device;
    c->a.step_limit = INT64_MAX;
    c->stream = static_cast<hipStream_t>(stream);
    DeviceGuard g(device);
    if (int rc = check(hipEventCreate(&c->ev0), "hipEventCreate")) { delete c; return rc; }
    if (int rc = check(hipEventCreate(&c->ev1), "hipEventCreate")) { delete c; return rc; }
    if (int rc = check(hipMalloc(&c->energy_partial, sizeof(double) * 2 * energy_partials()), "hipMalloc")) {
        fdtd_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int fdtd_ctx_destroy(fdtd_ctx* c) {
    if (!c) return 0;
    DeviceGuard g(c->device);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->mic_idx_dev) (void)hipFree(c->mic_idx_dev);
    if (c->mic_w_dev) (void)hipFree(c->mic_w_dev);
    if (c->mic_f64_dev) (void)hipFree(c->mic_f64_dev);
    for (fdtd_ctx::DevBuf* b : {&c->ade_cells, &c->m_cell, &c->m_estart, &c->m_esrc, &c->m_efield, &c->m_eweight, &c->m_pstart,
                                &c->m_pidx, &c->ade_poles, &c->probes})
        if (b->p) (void)hipFree(b->p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->sweep_ev) (void)hipEventDestroy(e);
    if (c->energy_partial) (void)hipFree(c->energy_partial);
    delete c;
    return 0;
}

static void drop_graph(fdtd_ctx* c) {
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        c->graph_steps = 0;
    }
}

int fdtd_ctx_set_grid(fdtd_ctx* c, int nx, int ny, int nz, int nonuniform, float cv, float cp,
                      const float* ifx, const float* ify, const float* ifz, const float* icx,
                      const float* icy, const float* icz) {
    if (!c) return fail(-1, "null context");
    if (nx < 2 || ny < 2 || nz < 2) return fail(-2, "grid %dx%dx%d: every axis needs >= 2 cells", nx, ny, nz);
    if (static_cast<int64_t>(nx) * ny * nz >= (int64_t(1) << 40)) return fail(-2, "grid too large");
    if (nonuniform && (!ifx || !ify || !ifz || !icx || !icy || !icz))
        return fail(-2, "nonuniform grid needs all six spacing tables");
    drop_graph(c);
    c->a.nx = nx; c->a.ny = ny; c->a.nz = nz;
    c->a.cv = cv; c->a.cp = cp;
    c->nonuniform = nonuniform != 0;
    c->a.ifc[0] = ifx; c->a.ifc[1] = ify; c->a.ifc[2] = ifz;
    c->a.icl[0] = icx; c->a.icl[1] = icy; c->a.icl[2] = icz;
    c->have_grid = true;
    return 0;
}

int fdtd_ctx_set_fields(fdtd_ctx* c, float* p0, float* vx0, float* vy0, float* vz0, float* p1,
                        float* vx1, float* vy1, float* vz1) {
    if (!c) return fail(-1, "null context");
    float* ptrs[8] = {p0, vx0, vy0, vz0, p1, vx1, vy1, vz1};
    for (auto* q : ptrs)
        if (!q) return fail(-2, "fdtd_ctx_set_fields: null field pointer");
    drop_graph(c);
    for (int s = 0; s < 2; ++s)
        for (int f = 0; f < 4; ++f) c->a.f[s][f] = ptrs[4 * s + f];
    c->have_fields = true;
    return 0;
}

int fdtd_ctx_set_flags(fdtd_ctx* c, const uint8_t* flags) {
    if (!c) return fail(-1, "null context");
    if (!flags) return fail(-2, "null flags");
    drop_graph(c);
    c->a.flags = flags;
    c->have_flags = true;
    return 0;
}

int fdtd_ctx_set_pml(fdtd_ctx* c, const float* dx, const float* dy, const float* dz) {
    if (!c) return fail(-1, "null context");
    if (!dx || !dy || !dz) return fail(-2, "PML tables must all be set (use 1.0f tables for an axis without PML)");
    drop_graph(c);
    c->a.dec[0] = dx; c->a.dec[1] = dy; c->a.dec[2] = dz;
    c->have_pml = true;
    return 0;
}

int fdtd_ctx_set_sources(fdtd_ctx* c, int n, const int32_t* kind, const int64_t* a, const int64_t* b,
                         const double* wave, int64_t wave_len) {
    if (!c) return fail(-1, "null context");
    if (!c->have_grid) return fail(-3, "set the grid before the sources");
    if (n < 0 || n > FDTD_MAX_SOURCES) return fail(-2, "%d sources (max %d)", n, FDTD_MAX_SOURCES);
    if (n > 0 && (!kind || !a || !b || !wave || wave_len < 1)) return fail(-2, "bad source arrays");
    const int nx = c->a.nx, ny = c->a.ny, nz = c->a.nz;
    for (int q = 0; q < n; ++q) {
        if (kind[q] == 0) {
            const int64_t f = a[q];
            if (f < 0 || f >= static_cast<int64_t>(nx) * ny * nz)
                return fail(-2, "point source %d: flat index %lld outside the grid", q, (long long)f);
            c->a.src_kind[q] = 0;
            c->a.src_k[q] = static_cast<int>(f % nz);
            c->a.src_j[q] = static_cast<int>((f / nz) % ny);
            c->a.src_i[q] = static_cast<int>(f / (static_cast<int64_t>(ny) * nz));
        } else if (kind[q] == 1) {
            const int ax = static_cast<int>(a[q]);
            const int dims[3] = {nx, ny, nz};
            if (ax < 0 || ax > 2 || b[q] < 0 || b[q] >= dims[ax])
                return fail(-2, "plane source %d: axis %d index %lld outside the grid", q, ax, (long long)b[q]);
            c->a.src_kind[q] = 1;
            c->a.src_i[q] = ax;
            c->a.src_j[q] = static_cast<int>(b[q]);
            c->a.src_k[q] = 0;
        } else {
            return fail(-2, "source %d: unknown kind %d", q, kind[q]);
        }
    }
    drop_graph(c);
    c->a.n_src = n;
    c->a.wave = wave;
    c->a.wave_len = n > 0 ? wave_len : 0;
    return 0;
}

static int upload(fdtd_ctx::DevBuf& b, const void* host, size_t bytes);

int fdtd_ctx_set_probes(fdtd_ctx* c, int n, const int64_t* idx, float* trace, int64_t trace_len) {
    if (!c) return fail(-1, "null context");
    if (!c->have_grid) return fail(-3, "set the grid before the probes");
    if (n < 0 || n > FDTD_MAX_PROBES) return fail(-2, "%d probes (max %d)", n, FDTD_MAX_PROBES);
    if (n > 0 && (!idx || !trace || trace_len < 1)) return fail(-2, "bad probe arrays");
    const int nx = c->a.nx, ny = c->a.ny, nz = c->a.nz;
    std::vector<int32_t> host(3 * static_cast<size_t>(std::max(n, 1)), 0);
    for (int q = 0; q < n; ++q) {
        const int64_t f = idx[q];
        if (f < 0 || f >= static_cast<int64_t>(nx) * ny * nz)
            return fail(-2, "probe %d: flat index %lld outside the grid", q, (long long)f);
        host[3 * q + 2] = static_cast<int32_t>(f % nz);
        host[3 * q + 1] = static_cast<int32_t>((f / nz) % ny);
        host[3 * q + 0] = static_cast<int32_t>(f / (static_cast<int64_t>(ny) * nz));
    }
    DeviceGuard g(c->device);
    drop_graph(c);
    // steps already enqueued on the context stream still read the old coordinates
    if (int rc = check(hipStreamSynchronize(c->stream), "stream sync")) return rc;
    if (int rc = upload(c->probes, host.data(), sizeof(int32_t) * host.size())) return rc;
    c->a.n_probe = n;
    // with boundary planes every probe is recorded after them
    c->a.n_sweep_probe = c->bnd.n ? 0 : std::min(n, kSweepProbes);
    c->a.probe_ijk = static_cast<const int32_t*>(c->probes.p);
    c->a.trace = trace;
    c->a.trace_len = n > 0 ? trace_len : 0;
    return 0;
}

int fdtd_ctx_set_microphones(fdtd_ctx* c, int n, const int32_t* idx, const double* w, const int32_t* f64,
                             double* out, int64_t out_len) {
    if (!c) return fail(-1, "null context");
    if (!c->have_grid) return fail(-3, "set the grid before the microphones");
    if (n < 0 || n > FDTD_MAX_MICS) return fail(-2, "%d microphones (max %d)", n, FDTD_MAX_MICS);
    if (n > 0 && (!idx || !w || !f64 || !out || out_len < 1)) return fail(-2, "bad microphone arrays");
    const int64_t cells = static_cast<int64_t>(c->a.nx) * c->a.ny * c->a.nz;
    for (int64_t q = 0; q < 32 * static_cast<int64_t>(n); ++q)
        if (idx[q] < 0 || idx[q] >= cells)
            return fail(-2, "microphone %lld: flat index %d outside the grid", (long long)(q / 32), idx[q]);
    DeviceGuard g(c->device);
    drop_graph(c);
    // steps already enqueued on the context stream still read the old tables
    if (int rc = check(hipStreamSynchronize(c->stream), "stream sync")) return rc;
    if (n > c->mic_cap) {
        // set-up allocation (never on a step path): the gather tables
        if (c->mic_idx_dev) (void)hipFree(c->mic_idx_dev);
        if (c->mic_w_dev) (void)hipFree(c->mic_w_dev);
        if (c->mic_f64_dev) (void)hipFree(c->mic_f64_dev);
        c->mic_idx_dev = nullptr;
        c->mic_w_dev = nullptr;
        c->mic_f64_dev = nullptr;
        c->mic_cap = 0;
        if (int rc = check(hipMalloc(&c->mic_idx_dev, sizeof(int32_t) * 32 * n), "hipMalloc")) return rc;
        if (int rc = check(hipMalloc(&c->mic_w_dev, sizeof(double) * 32 * n), "hipMalloc")) return rc;
        if (int rc = check(hipMalloc(&c->mic_f64_dev, sizeof(int32_t) * 4 * n), "hipMalloc")) return rc;
        c->mic_cap = n;
    }
    if (n > 0) {
        if (int rc = check(hipMemcpy(c->mic_idx_dev, idx, sizeof(int32_t) * 32 * n, hipMemcpyHostToDevice),
                           "hipMemcpy"))
            return rc;
        if (int rc = check(hipMemcpy(c->mic_w_dev, w, sizeof(double) * 32 * n, hipMemcpyHostToDevice),
                           "hipMemcpy"))
            return rc;
        if (int rc = check(hipMemcpy(c->mic_f64_dev, f64, sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice),
                           "hipMemcpy"))
            return rc;
    }
    c->a.n_mic = n;
    c->a.mic_idx = c->mic_idx_dev;
    c->a.mic_w = c->mic_w_dev;
    c->a.mic_f64 = c->mic_f64_dev;
    c->a.mic_out = out;
    c->a.mic_len = n > 0 ? out_len : 0;
    return 0;
}

// Copy a host table into an owned device buffer (set-up only: grows with
// hipMalloc when needed; the copy is synchronous).
static int upload(fdtd_ctx::DevBuf& b, const void* host, size_t bytes) {
    if (bytes > b.bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (int rc = check(hipMalloc(&b.p, bytes), "hipMalloc")) return rc;
        b.bytes = bytes;
    }
    if (bytes) return check(hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return 0;
}

int fdtd_ctx_set_membranes(fdtd_ctx* c, int n_cells, const int64_t* cell, const int32_t* entry_start,
                           int n_entries, const int32_t* entry_src, const int32_t* entry_field,
                           const double* entry_weight, const int32_t* probe_start, int n_probe_refs,
                           const int32_t* probe_idx, int n_src, const double* wave, int64_t wave_len) {
    if (!c) return fail(-1, "null context");
    if (!c->have_grid) return fail(-3, "set the grid before the membrane sources");
    if (n_cells < 0 || n_entries < 0 || n_probe_refs < 0 || n_src < 0) return fail(-2, "negative table size");
    if (n_cells > 0) {
        if (!cell || !entry_start || !entry_src || !entry_field || !entry_weight || !probe_start ||
            (n_probe_refs > 0 && !probe_idx) || !wave || wave_len < 1 || n_src < 1)
            return fail(-2, "bad membrane arrays");
        const int64_t cells = static_cast<int64_t>(c->a.nx) * c->a.ny * c->a.nz;
        if (entry_start[0] != 0 || entry_start[n_cells] != n_entries || probe_start[0] != 0 ||
            probe_start[n_cells] != n_probe_refs)
            return fail(-2, "membrane CSR tables do not span their entries");
        for (int e = 0; e < n_cells; ++e) {
            if (cell[e] < 0 || cell[e] >= cells || cell[e] > INT32_MAX)
                return fail(-2, "source table cell %d: flat index %lld outside the grid or above 2^31-1", e,
                            (long long)cell[e]);
            if (e > 0 && cell[e] <= cell[e - 1]) return fail(-2, "membrane cells must be increasing");
            if (entry_start[e + 1] <= entry_start[e] || probe_start[e + 1] < probe_start[e])
                return fail(-2, "membrane cell %d: bad entry or probe range", e);
        }
        for (int t = 0; t < n_entries; ++t)
            if (entry_src[t] < 0 || entry_src[t] >= n_src || entry_field[t] < 0 || entry_field[t] > 3)
                return fail(-2, "membrane entry %d: source %d / field %d out of range", t, entry_src[t],
                            entry_field[t]);
        for (int t = 0; t < n_probe_refs; ++t)
            if (probe_idx[t] < 0 || probe_idx[t] >= c->a.n_probe)
                return fail(-2, "membrane probe reference %d: probe %d not set", t, probe_idx[t]);
    }
    DeviceGuard g(c->device);
    drop_graph(c);
    // steps already on the stream may still read the old tables
    if (int rc = check(hipStreamSynchronize(c->stream), "stream sync")) return rc;
    c->a.n_mcell = 0;
    if (n_cells == 0) return 0;
    std::vector<int32_t> cell32(cell, cell + n_cells);
    if (int rc = upload(c->m_cell, cell32.data(), sizeof(int32_t) * n_cells)) return rc;
    if (int rc = upload(c->m_estart, entry_start, sizeof(int32_t) * (n_cells + 1))) return rc;
    if (int rc = upload(c->m_esrc, entry_src, sizeof(int32_t) * n_entries)) return rc;
    if (int rc = upload(c->m_efield, entry_field, sizeof(int32_t) * n_entries)) return rc;
    if (int rc = upload(c->m_eweight, entry_weight, sizeof(double) * n_entries)) return rc;
    if (int rc = upload(c->m_pstart, probe_start, sizeof(int32_t) * (n_cells + 1))) return rc;
    if (int rc = upload(c->m_pidx, probe_idx, sizeof(int32_t) * n_probe_refs)) return rc;
    c->a.mcell = static_cast<const int32_t*>(c->m_cell.p);
    c->a.m_estart = static_cast<const int32_t*>(c->m_estart.p);
    c->a.m_esrc = static_cast<const int32_t*>(c->m_esrc.p);
    c->a.m_efield = static_cast<const int32_t*>(c->m_efield.p);
    c->a.m_eweight = static_cast<const double*>(c->m_eweight.p);
    c->a.m_pstart = static_cast<const int32_t*>(c->m_pstart.p);
    c->a.m_pidx = static_cast<const int32_t*>(c->m_pidx.p);
    c->a.m_wave = wave;
    c->a.m_wave_len = wave_len;
    c->a.m_single = n_entries == n_cells ? 1 : 0;  // every cell has at least one entry
    c->a.n_mcell = n_cells;
    return 0;
}

int fdtd_ctx_set_boundary_planes(fdtd_ctx* c, int n, const fdtd_boundary_plane* planes, float* state,
                                 int64_t state_len) {
    if (!c) return fail(-1, "null context");
    if (n < 0 || n > FDTD_MAX_BOUNDARY_PLANES) return fail(-2, "%d boundary planes (max %d)", n, FDTD_MAX_BOUNDARY_PLANES);
    if (n > 0 && (!planes || !state)) return fail(-2, "null boundary arrays");
    if (n > 0 && !c->have_grid) return fail(-3, "set the grid before the boundary planes");
    const int dims[3] = {c->a.nx, c->a.ny, c->a.nz};
    BoundarySet b;
    for (int q = 0; q < n; ++q) {
        const fdtd_boundary_plane& p = planes[q];
        if (p.kind != 0 && p.kind != 1) return fail(-2, "boundary plane %d: kind %d", q, p.kind);
        if (p.axis < 0 || p.axis > 2) return fail(-2, "boundary plane %d: axis %d", q, p.axis);
        if (p.kind == 1 && p.side != 0 && p.side != 1) return fail(-2, "boundary plane %d: side %d", q, p.side);
        if (dims[p.axis] < 2) return fail(-2, "boundary plane %d: its axis has %d cell(s), needs 2", q, dims[p.axis]);
        const int64_t cross = static_cast<int64_t>(dims[0]) * dims[1] * dims[2] / dims[p.axis];
        const int64_t need = cross * (p.kind == 0 ? 2 : 1);
        if (p.state_off < 0 || p.state_off + need > state_len)
            return fail(-2, "boundary plane %d: state [%lld, %lld) outside the %lld-element state", q,
                        (long long)p.state_off, (long long)(p.state_off + need), (long long)state_len);
        b.planes[q] = p;
    }
    b.n = n;
    b.state = n ? state : nullptr;
    drop_graph(c);
    c->bnd = b;
    c->a.n_sweep_probe = n ? 0 : std::min(c->a.n_probe, kSweepProbes);
    return 0;
}

int fdtd_ctx_set_counters(fdtd_ctx* c, int64_t* step_counter, uint32_t* health) {
    if (!c) return fail(-1, "null context");
    if (!step_counter || !health) return fail(-2, "null counter");
    drop_graph(c);
    c->a.step_ptr = step_counter;
    c->step_counter = step_counter;
    c->a.health = health;
    c->have_counters = true;
    return 0;
}

int fdtd_ctx_set_ade(fdtd_ctx* c, int n_pole, const int32_t* mat_id, const int32_t* target,
                     const int32_t* lorentz, const float* pa, const float* pb, const float* pd,
                     const float* corr, const uint8_t* material, const int32_t* lo, const int32_t* hi,
                     float* J, float* Jprev, float inv_dx) {
    if (!c) return fail(-1, "null context");
    if (n_pole < 0 || n_pole > FDTD_MAX_POLES) return fail(-2, "%d poles (max %d)", n_pole, FDTD_MAX_POLES);
    drop_graph(c);
    // Everything is checked and built in `a` first; the context changes only
    // when the call succeeds, and a failed call leaves it without poles (never
    // with a pole count over stale tables).
    c->a.n_pole = 0;
    if (n_pole == 0) return 0;
    if (!mat_id || !target || !lorentz || !pa || !pb || !pd || !corr || !material || !lo || !hi || !J)
        return fail(-2, "bad ADE arrays");
    if (!c->have_grid) return fail(-3, "set the grid before the ADE materials");
    FusedArgs a = c->a;
    const int dims[3] = {a.nx, a.ny, a.nz};
    for (int ax = 0; ax < 3; ++ax) {
        if (lo[ax] < 0 || hi[ax] > dims[ax] || lo[ax] >= hi[ax])
            return fail(-2, "ADE box axis %d [%d,%d) outside the grid", ax, lo[ax], hi[ax]);
        a.box_lo[ax] = lo[ax];
        a.box_hi[ax] = hi[ax];
    }
    std::vector<AdePole> poles(n_pole);
    for (int q = 0; q < n_pole; ++q) {
        if (mat_id[q] < 1 || mat_id[q] > 255) return fail(-2, "pole %d: material id %d", q, mat_id[q]);
        if (target[q] != 0 && target[q] != 1) return fail(-2, "pole %d: target %d", q, target[q]);
        if (lorentz[q] != 0 && lorentz[q] != 1) return fail(-2, "pole %d: kind %d", q, lorentz[q]);
        if (lorentz[q] && !Jprev) return fail(-2, "Lorentz pole without a history array");
        poles[q] = AdePole{mat_id[q], target[q], lorentz[q], 0, pa[q], pb[q], pd[q], corr[q]};
    }
    a.material = material;
    a.J = J;
    a.Jprev = Jprev;
    a.ade_inv_dx = inv_dx;
    // The box cells whose material has poles, as box indices in box order: the
    // ADE kernels then run one thread per such cell rather than per box cell
    // (a sphere fills about half its box). Read from the device material ids
    // one box plane at a time, after the work already on the stream.
    a.ade_cells = nullptr;
    a.n_ade_cells = 0;
    const int bx = hi[0] - lo[0], by = hi[1] - lo[1], bz = hi[2] - lo[2];
    const int64_t box_n = static_cast<int64_t>(bx) * by * bz;
    std::vector<int32_t> list;
    DeviceGuard g(c->device);
    // the material ids may come from any stream; steps in flight read the old tables
    if (int rc = check(hipDeviceSynchronize(), "device sync")) return rc;
    if (box_n < INT32_MAX) {
        bool has_pole[256] = {};
        for (int q = 0; q < n_pole; ++q) has_pole[mat_id[q]] = true;
        const size_t plane = static_cast<size_t>(by) * a.nz;
        std::vector<uint8_t> mat(plane);
        for (int i = lo[0]; i < hi[0]; ++i) {
            const uint8_t* src = material + (static_cast<int64_t>(i) * a.ny + lo[1]) * a.nz;
            if (int rc = check(hipMemcpy(mat.data(), src, plane, hipMemcpyDeviceToHost), "hipMemcpy")) return rc;
            for (int j = 0; j < by; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    const uint8_t m = mat[static_cast<size_t>(j) * a.nz + k];
                    if (m && has_pole[m])
                        list.push_back(static_cast<int32_t>(((static_cast<int64_t>(i - lo[0]) * by) + j) * bz +
                                                            (k - lo[2])));
                }
        }
        const int32_t none = 0;
        if (int rc = upload(c->ade_cells, list.empty() ? &none : list.data(),
                            sizeof(int32_t) * std::max<size_t>(list.size(), 1)))
            return rc;
        a.ade_cells = static_cast<const int32_t*>(c->ade_cells.p);
        a.n_ade_cells = static_cast<int64_t>(list.size());
    }
    if (int rc = upload(c->ade_poles, poles.data(), sizeof(AdePole) * poles.size())) return rc;
    a.poles = static_cast<const AdePole*>(c->ade_poles.p);
    a.n_pole = n_pole;
    c->a = a;
    return 0;
}

int fdtd_ctx_set_tiling(fdtd_ctx* c, int rows, int planes, int stack) {
    if (!c) return fail(-1, "null context");
    if (rows == 0) rows = 4;
    if (planes == 0) planes = 8;
    if (stack == 0) stack = 1;
    if (rows != 2 && rows != 4 && rows != 8 && rows != 16) return fail(-2, "rows must be 2, 4, 8 or 16");
    if (planes < 1) return fail(-2, "planes must be >= 1");
    if (stack != 1 && stack != 2 && stack != 4 && stack != 8) return fail(-2, "stack must be 1, 2, 4 or 8");
    drop_graph(c);
    c->rows = rows;
    c->planes = planes;
    c->stack = stack;
    return 0;
}

static int ready(fdtd_ctx* c) {
    if (!c) return fail(-1, "null context");
    if (!c->have_grid || !c->have_fields || !c->have_flags || !c->have_pml || !c->have_counters)
        return fail(-3, "context incomplete: grid=%d fields=%d flags=%d pml=%d counters=%d", c->have_grid,
                    c->have_fields, c->have_flags, c->have_pml, c->have_counters);
    if (c->stack > 1 && static_cast<int64_t>(c->a.nx) * c->a.ny * c->a.nz >= kMaxCells32)
        return fail(-2, "stacked sweeps need fewer than 2^30 cells; use stack 1");
    if (c->bnd.n && c->a.n_src)
        return fail(-2, "with boundary planes every source goes in the per-cell tables (the sweep adds %d)",
                    c->a.n_src);
    return 0;
}

static int enqueue_steps(fdtd_ctx* c, int n) {
    FusedArgs a = c->a;
    const int blocks = fused_grid(a, c->rows, c->planes, c->stack);
    for (int s = 0; s < n; ++s) {
        a.koff = s;
        if (int rc = check(launch_fused(a, c->rows, c->stack, c->nonuniform, a.n_pole > 0, blocks, c->stream,
                                        nullptr, nullptr, c->bnd.n ? &c->bnd : nullptr),
                           "fused step launch"))
            return rc;
    }
    return check(launch_advance(c->step_counter, n, c->stream), "advance launch");
}

int fdtd_step_n(fdtd_ctx* c, int n_steps) {
    if (int rc = ready(c)) return rc;
    if (n_steps < 0) return fail(-2, "negative step count");
    if (n_steps == 0) return 0;
    DeviceGuard g(c->device);
    return enqueue_steps(c, n_steps);
}

int fdtd_graph_run(fdtd_ctx* c, int n_steps, int replays) {
    if (int rc = ready(c)) return rc;
    if (n_steps < 1 || replays < 0) return fail(-2, "bad graph size");
    DeviceGuard g(c->device);
    if (!c->graph || c->graph_steps != n_steps) {
        drop_graph(c);
        hipStream_t cap = c->stream;
        bool own = false;
        if (!cap) {
            // the legacy default stream cannot be captured: capture on a private one
            if (int rc = check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate")) return rc;
            own = true;
        }
        hipGraph_t gr = nullptr;
        if (int rc = check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture")) {
            if (own) (void)hipStreamDestroy(cap);
            return rc;
        }
        hipStream_t saved = c->stream;
        c->stream = cap;
        int rc = enqueue_steps(c, n_steps);
        c->stream = saved;
        hipError_t e = hipStreamEndCapture(cap, &gr);
        if (own) (void)hipStreamDestroy(cap);
        if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
        if (int r2 = check(e, "end capture")) return r2;
        e = hipGraphInstantiate(&c->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (int r3 = check(e, "graph instantiate")) return r3;
        c->graph_steps = n_steps;
    }
    for (int r = 0; r < replays; ++r)
        if (int rc = check(hipGraphLaunch(c->graph, c->stream), "graph launch")) return rc;
    return 0;
}

int fdtd_time_steps(fdtd_ctx* c, int n_steps, int use_graph, float* ms_total) {
    if (int rc = ready(c)) return rc;
    if (!ms_total) return fail(-2, "null output");
    DeviceGuard g(c->device);
    if (int rc = check(hipEventRecord(c->ev0, c->stream), "event record")) return rc;
    int rc = use_graph ? fdtd_graph_run(c, n_steps, 1) : fdtd_step_n(c, n_steps);
    if (rc) return rc;
    if (int r2 = check(hipEventRecord(c->ev1, c->stream), "event record")) return r2;
    if (int r3 = check(hipEventSynchronize(c->ev1), "event sync")) return r3;
    return check(hipEventElapsedTime(ms_total, c->ev0, c->ev1), "elapsed time");
}

int fdtd_stream_ceiling(fdtd_ctx* c, int reps, float* ms_per_pass) {
    if (int rc = ready(c)) return rc;
    if (!ms_per_pass) return fail(-2, "null output");
    if (reps < 1 || reps > 1000) return fail(-2, "reps must be in [1, 1000]");
    for (int set = 0; set < 2; ++set)
        for (int g = 0; g < 4; ++g)
            if (reinterpret_cast<uintptr_t>(c->a.f[set][g]) % 16)
                return fail(-2, "field buffers must be 16-byte aligned for the streaming kernel");
    if (reinterpret_cast<uintptr_t>(c->a.flags) % 4) return fail(-2, "flag buffer must be 4-byte aligned");
    DeviceGuard g(c->device);
    if (int rc = check(launch_stream_ceiling(c->a, c->stream), "stream launch")) return rc;  // warm-up
    if (int rc = check(hipEventRecord(c->ev0, c->stream), "event record")) return rc;
    for (int r = 0; r < reps; ++r)
        if (int rc = check(launch_stream_ceiling(c->a, c->stream), "stream launch")) return rc;
    if (int rc = check(hipEventRecord(c->ev1, c->stream), "event record")) return rc;
    if (int rc = check(hipEventSynchronize(c->ev1), "event sync")) return rc;
    float ms = 0.f;
    if (int rc = check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed time")) return rc;
    *ms_per_pass = ms / reps;
    return 0;
}

int fdtd_stream_ceiling_march(fdtd_ctx* c, int reps, float* ms_per_pass) {
    if (int rc = ready(c)) return rc;
    if (!ms_per_pass) return fail(-2, "null output");
    if (reps < 1 || reps > 1000) return fail(-2, "reps must be in [1, 1000]");
    if (static_cast<int64_t>(c->a.nx) * c->a.ny * c->a.nz >= kMaxCells32)
        return fail(-2, "the marching pass takes grids below 2^30 cells");
    if (c->rows != 2 && c->rows != 4 && c->rows != 8) return fail(-2, "rows must be 2, 4 or 8");
    DeviceGuard g(c->device);
    auto pass = [&]() { return launch_stream_march(c->a, c->rows, c->planes, c->stream); };
    if (int rc = check(pass(), "stream launch")) return rc;  // warm-up
    if (int rc = check(hipEventRecord(c->ev0, c->stream), "event record")) return rc;
    for (int r = 0; r < reps; ++r)
        if (int rc = check(pass(), "stream launch")) return rc;
    if (int rc = check(hipEventRecord(c->ev1, c->stream), "event record")) return rc;
    if (int rc = check(hipEventSynchronize(c->ev1), "event sync")) return rc;
    float ms = 0.f;
    if (int rc = check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed time")) return rc;
    *ms_per_pass = ms / reps;
    return 0;
}

int fdtd_ctx_energy(fdtd_ctx* c, double* out) {
    if (int rc = ready(c)) return rc;
    if (!out) return fail(-2, "null output");
    for (int set = 0; set < 2; ++set)
        for (int g = 0; g < 4; ++g)
            if (reinterpret_cast<uintptr_t>(c->a.f[set][g]) % 16)
                return fail(-2, "field buffers must be 16-byte aligned for the energy sums");
    if (reinterpret_cast<uintptr_t>(c->a.flags) % 4) return fail(-2, "flag buffer must be 4-byte aligned");
    DeviceGuard g(c->device);
    return check(launch_energy(c->a, c->energy_partial, out, c->stream), "energy launch");
}

int fdtd_ctx_centred_velocity(fdtd_ctx* c, float* vx, float* vy, float* vz) {
    if (int rc = ready(c)) return rc;
    if (!vx || !vy || !vz) return fail(-2, "null output");
    DeviceGuard g(c->device);
    return check(launch_centre_velocity(c->a, vx, vy, vz, c->stream), "centring launch");
}

int fdtd_time_sweeps(fdtd_ctx* c, int n_steps, float* ms_sweep, float* ms_total) {
    if (int rc = ready(c)) return rc;
    if (!ms_sweep || !ms_total) return fail(-2, "null output");
    if (n_steps < 1 || n_steps > 4096) return fail(-2, "n_steps must be in [1, 4096]");
    DeviceGuard g(c->device);
    while (c->sweep_ev.size() < 2 * static_cast<size_t>(n_steps)) {
        hipEvent_t e = nullptr;
        if (int rc = check(hipEventCreate(&e), "event create")) return rc;
        c->sweep_ev.push_back(e);
    }
    FusedArgs a = c->a;
    const int blocks = fused_grid(a, c->rows, c->planes, c->stack);
    if (int rc = check(hipEventRecord(c->ev0, c->stream), "event record")) return rc;
    for (int s = 0; s < n_steps; ++s) {
        a.koff = s;
        if (int rc = check(launch_fused(a, c->rows, c->stack, c->nonuniform, a.n_pole > 0, blocks, c->stream,
                                        c->sweep_ev[2 * s], c->sweep_ev[2 * s + 1], c->bnd.n ? &c->bnd : nullptr),
                           "fused step launch"))
            return rc;
    }
    if (int rc = check(launch_advance(c->step_counter, n_steps, c->stream), "advance launch")) return rc;
    if (int rc = check(hipEventRecord(c->ev1, c->stream), "event record")) return rc;
    if (int rc = check(hipEventSynchronize(c->ev1), "event sync")) return rc;
    float total = 0.f, sweep = 0.f;
    if (int rc = check(hipEventElapsedTime(&total, c->ev0, c->ev1), "elapsed time")) return rc;
    for (int s = 0; s < n_steps; ++s) {
        float ms = 0.f;
        if (int rc = check(hipEventElapsedTime(&ms, c->sweep_ev[2 * s], c->sweep_ev[2 * s + 1]), "elapsed time"))
            return rc;
        sweep += ms;
    }
    *ms_sweep = sweep;
    *ms_total = total;
    return 0;
}

int64_t fdtd_voxelize_scratch_bytes(int n_ops, int n_prog, int n_params) {
    if (n_ops < 0 || n_prog < 1 || n_params < 0) return -1;
    const int64_t ints = 2 * static_cast<int64_t>(n_ops) + (n_prog + 1) + n_prog;
    return ((ints * 4 + 7) / 8) * 8 + 8 * static_cast<int64_t>(n_params);
}

}  // extern "C"

namespace {
// Validates a voxeliser call, stages its programs into `base` (device memory
// of `bytes` bytes) on `st`, waits for the upload (the host arrays belong to
// the caller) and launches the kernel. `who` prefixes error messages.
int voxelize_staged(const char* who, const int32_t* ops, int n_ops, const int32_t* prog_start,
                    const int32_t* prog_ids, int n_prog, const double* params, int n_params, void* base_v,
                    int64_t bytes, const double* xc, const double* yc, const double* zc, int nx, int ny, int nz,
                    uint8_t* air, uint8_t* ids, hipStream_t st) {
    if (!prog_start || !prog_ids || (n_ops > 0 && !ops) || (n_params > 0 && !params))
        return fail(-1, "%s: null program array", who);
    if (!base_v || !xc || !yc || !zc || !air) return fail(-1, "%s: null device pointer", who);
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "%s: bad dimensions", who);
    if (n_ops < 0 || n_ops > (1 << 20) || n_params < 0 || n_params > (1 << 22))
        return fail(-2, "%s: program too large", who);
    if (n_prog > 1 && !ids) return fail(-1, "%s: material programs need an ids output", who);
    if (const char* msg = vox_validate(ops, n_ops, prog_start, n_prog, prog_ids, params, n_params))
        return fail(-2, "%s: %s", who, msg);
    const int64_t need = fdtd_voxelize_scratch_bytes(n_ops, n_prog, n_params);
    if (bytes < need)
        return fail(-2, "%s: programs need %lld B of staging, %lld B available", who, (long long)need,
                    (long long)bytes);
    auto* base = static_cast<char*>(base_v);
    int32_t* d_ops = reinterpret_cast<int32_t*>(base);
    int32_t* d_start = d_ops + 2 * static_cast<int64_t>(n_ops);
    int32_t* d_ids = d_start + (n_prog + 1);
    const int64_t ints = 2 * static_cast<int64_t>(n_ops) + (n_prog + 1) + n_prog;
    double* d_params = reinterpret_cast<double*>(base + ((ints * 4 + 7) / 8) * 8);
    if (n_ops)
        if (int rc = check(hipMemcpyAsync(d_ops, ops, 8 * static_cast<size_t>(n_ops), hipMemcpyHostToDevice, st),
                           "voxelize program upload"))
            return rc;
    if (int rc = check(hipMemcpyAsync(d_start, prog_start, 4 * static_cast<size_t>(n_prog + 1),
                                      hipMemcpyHostToDevice, st), "voxelize program upload"))
        return rc;
    if (int rc = check(hipMemcpyAsync(d_ids, prog_ids, 4 * static_cast<size_t>(n_prog), hipMemcpyHostToDevice, st),
                       "voxelize program upload"))
        return rc;
    if (n_params)
        if (int rc = check(hipMemcpyAsync(d_params, params, 8 * static_cast<size_t>(n_params),
                                          hipMemcpyHostToDevice, st), "voxelize program upload"))
            return rc;
    // the host arrays belong to the caller: finish reading them before returning
    if (int rc = check(hipStreamSynchronize(st), "voxelize upload sync")) return rc;
    VoxProgram pg{d_ops, d_start, d_ids, d_params, n_prog};
    return check(launch_voxelize(pg, xc, yc, zc, nx, ny, nz, air, ids, st), "voxelize launch");
}

// Program staging of fdtd_voxelize_static: a fixed array in device memory, one
// copy per GPU (each GPU holds it until the process ends). In host code the
// name `scratch` is only a host-side stand-in; every use goes through the
// address hipGetSymbolAddress returns for the GPU (scratch_addr[device]).
__device__ __align__(256) static unsigned char scratch[FDTD_VOXELIZE_STATIC_BYTES];
constexpr int kMaxDevices = 64;
void* scratch_addr[kMaxDevices] = {};
std::mutex scratch_mu;

// Address of this GPU's copy of `scratch`, looked up once per device with the
// device current.
int static_scratch(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(-2, "fdtd_voxelize_static: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_addr[dev]) {
        DeviceGuard g(dev);
        void* p = nullptr;
        if (int rc = check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "fdtd_voxelize_static: scratch address"))
            return rc;
        if (!p) return fail(-5, "fdtd_voxelize_static: scratch address is null");
        scratch_addr[dev] = p;
    }
    *out = scratch_addr[dev];
    return 0;
}
}  // namespace

extern "C" {

int fdtd_voxelize(const int32_t* ops, int n_ops, const int32_t* prog_start, const int32_t* prog_ids, int n_prog,
                  const double* params, int n_params, void* dev_scratch, int64_t scratch_bytes,
                  const double* xc, const double* yc, const double* zc, int nx, int ny, int nz,
                  uint8_t* air, uint8_t* ids, void* stream) {
    return voxelize_staged("fdtd_voxelize", ops, n_ops, prog_start, prog_ids, n_prog, params, n_params,
                           dev_scratch, scratch_bytes, xc, yc, zc, nx, ny, nz, air, ids,
                           static_cast<hipStream_t>(stream));
}

int fdtd_voxelize_static(const int32_t* ops, int n_ops, const int32_t* prog_start, const int32_t* prog_ids,
                         int n_prog, const double* params, int n_params, const double* xc, const double* yc,
                         const double* zc, int nx, int ny, int nz, uint8_t* air, uint8_t* ids, void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = -1;
    if (st) {
        if (int rc = check(hipStreamGetDevice(st, &dev), "fdtd_voxelize_static: stream device")) return rc;
    } else if (int rc = check(hipGetDevice(&dev), "fdtd_voxelize_static: current device")) {
        return rc;
    }
    void* base = nullptr;
    if (int rc = static_scratch(dev, &base)) return rc;
    DeviceGuard g(dev);
    return voxelize_staged("fdtd_voxelize_static", ops, n_ops, prog_start, prog_ids, n_prog, params, n_params,
                           base, FDTD_VOXELIZE_STATIC_BYTES, xc, yc, zc, nx, ny, nz, air, ids, st);
}

// ---- per-phase entry points ---------------------------------------------------
#define CHECK_DIMS()                                                   \
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "bad dimensions")

int fdtd_update_velocity(const float* p, float* vx, float* vy, float* vz, int nx, int ny, int nz,
                         float cv, void* stream) {
    CHECK_DIMS();
    return check(phase_velocity(p, vx, vy, vz, nx, ny, nz, cv, nullptr, nullptr, nullptr,
                                static_cast<hipStream_t>(stream)), "update_velocity");
}

int fdtd_update_velocity_nonuniform(const float* p, float* vx, float* vy, float* vz, int nx, int ny,
                                    int nz, float cv_base, const float* ifx, const float* ify,
                                    const float* ifz, void* stream) {
    CHECK_DIMS();
    if (!ifx || !ify || !ifz) return fail(-2, "null spacing table");
    return check(phase_velocity(p, vx, vy, vz, nx, ny, nz, cv_base, ifx, ify, ifz,
                                static_cast<hipStream_t>(stream)), "update_velocity_nonuniform");
}

int fdtd_apply_rigid(float* vx, float* vy, float* vz, const uint8_t* flags, int nx, int ny, int nz,
                     void* stream) {
    CHECK_DIMS();
    return check(phase_rigid(vx, vy, vz, flags, nx, ny, nz, static_cast<hipStream_t>(stream)),
                 "apply_rigid");
}

int fdtd_update_pressure(float* p, const float* vx, const float* vy, const float* vz,
                         const uint8_t* flags, int nx, int ny, int nz, float cp, void* stream) {
    CHECK_DIMS();
    return check(phase_pressure(p, vx, vy, vz, flags, nx, ny, nz, cp, nullptr, nullptr, nullptr,
                                static_cast<hipStream_t>(stream)), "update_pressure");
}

int fdtd_update_pressure_nonuniform(float* p, const float* vx, const float* vy, const float* vz,
                                    const uint8_t* flags, int nx, int ny, int nz, float cp_base,
                                    const float* icx, const float* icy, const float* icz,
                                    void* stream) {
    CHECK_DIMS();
    if (!icx || !icy || !icz) return fail(-2, "null spacing table");
    return check(phase_pressure(p, vx, vy, vz, flags, nx, ny, nz, cp_base, icx, icy, icz,
                                static_cast<hipStream_t>(stream)), "update_pressure_nonuniform");
}

int fdtd_apply_pml_velocity(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                            const float* dz, int nx, int ny, int nz, void* stream) {
    CHECK_DIMS();
    return check(phase_pml_velocity(vx, vy, vz, dx, dy, dz, nx, ny, nz, static_cast<hipStream_t>(stream)),
                 "apply_pml_velocity");
}

int fdtd_apply_pml_pressure(float* p, const float* dx, const float* dy, const float* dz, int nx,
                            int ny, int nz, void* stream) {
    CHECK_DIMS();
    return check(phase_pml_pressure(p, dx, dy, dz, nx, ny, nz, static_cast<hipStream_t>(stream)),
                 "apply_pml_pressure");
}

int fdtd_compute_divergence(float* div, const float* vx, const float* vy, const float* vz, int nx,
                            int ny, int nz, float inv_dx, const float* icx, const float* icy,
                            const float* icz, void* stream) {
    CHECK_DIMS();
    return check(phase_divergence(div, vx, vy, vz, nx, ny, nz, inv_dx, icx, icy, icz,
                                  static_cast<hipStream_t>(stream)), "compute_divergence");
}

int fdtd_build_flags(const uint8_t* air, const uint8_t* material, uint8_t* flags, int nx, int ny,
                     int nz, void* stream) {
    CHECK_DIMS();
    if (!air || !flags) return fail(-2, "null mask");
    return check(phase_build_flags(air, material, flags, nx, ny, nz, static_cast<hipStream_t>(stream)),
                 "build_flags");
}

// ---- batched serving ----------------------------------------------------------
struct fdtd_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int rows = 4;
    int cap = 0;                        // solves the buffers hold
    FusedArgs* table_dev = nullptr;     // [cap]
    int* meta_dev = nullptr;            // [cap + 2]: n, then offsets[0..n]
    FusedArgs* table_host[2] = {};      // pinned staging, used alternately
    int* meta_host[2] = {};
    hipEvent_t copied[2] = {};          // the staging buffer's last upload has run
    bool pending[2] = {};
    int stage = 0;
    int n = 0;
    int total_waves = 0;
    bool nts = true;                    // non-temporal stores (the solves exceed the cache)
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int graph_waves = 0;                // wave capacity of the captured grid
    bool graph_nts = true;
};

static void batch_drop_graph(fdtd_batch* b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        b->graph_steps = 0;
        b->graph_waves = 0;
    }
}

static void batch_free(fdtd_batch* b) {
    if (b->table_dev) (void)hipFree(b->table_dev);
    if (b->meta_dev) (void)hipFree(b->meta_dev);
    for (int i = 0; i < 2; ++i) {
        if (b->table_host[i]) (void)hipHostFree(b->table_host[i]);
        if (b->meta_host[i]) (void)hipHostFree(b->meta_host[i]);
        b->table_host[i] = nullptr;
        b->meta_host[i] = nullptr;
        b->pending[i] = false;
    }
    b->table_dev = nullptr;
    b->meta_dev = nullptr;
    b->cap = 0;
}

int fdtd_batch_create(int device, void* stream, int rows, fdtd_batch** out) {
    if (!out) return fail(-1, "null out");
    if (rows != 2 && rows != 4 && rows != 8) return fail(-2, "batch rows must be 2, 4 or 8");
    DeviceGuard g(device);
    auto* b = new fdtd_batch();
    b->device = device;
    b->stream = static_cast<hipStream_t>(stream);
    b->rows = rows;
    for (int i = 0; i < 2; ++i)
        if (int rc = check(hipEventCreateWithFlags(&b->copied[i], hipEventDisableTiming), "hipEventCreate")) {
            delete b;
            return rc;
        }
    *out = b;
    return 0;
}

int fdtd_batch_destroy(fdtd_batch* b) {
    if (!b) return 0;
    DeviceGuard g(b->device);
    batch_drop_graph(b);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    else (void)hipDeviceSynchronize();
    batch_free(b);
    for (int i = 0; i < 2; ++i)
        if (b->copied[i]) (void)hipEventDestroy(b->copied[i]);
    delete b;
    return 0;
}

int fdtd_batch_set(fdtd_batch* b, fdtd_ctx* const* ctxs, const int64_t* step_limits, int n) {
    if (!b) return fail(-1, "null batch");
    if (n < 0 || (n > 0 && (!ctxs || !step_limits))) return fail(-2, "bad batch arrays");
    DeviceGuard g(b->device);
    for (int q = 0; q < n; ++q) {
        fdtd_ctx* c = ctxs[q];
        if (int rc = ready(c)) return rc;
        if (c->a.n_probe > kSweepProbes)
            return fail(-2, "batch entry %d: %d probes (batched solves record at most %d)", q, c->a.n_probe,
                        kSweepProbes);
        if (c->nonuniform || c->a.n_pole > 0 || c->a.n_mic > 0 || c->a.n_mcell > 0 || c->bnd.n > 0)
            return fail(-2, "batch entry %d: nonuniform, ADE, microphone, membrane-source or boundary-plane "
                            "solves run on their own context", q);
        if (c->device != b->device) return fail(-2, "batch entry %d is on another device", q);
        if (static_cast<int64_t>(c->a.nx) * c->a.ny * c->a.nz >= kMaxCells32)
            return fail(-2, "batch entry %d: batched solves need fewer than 2^30 cells", q);
    }
    if (n > b->cap || !b->meta_dev) {
        // growth (rare): no queued work may still read the old buffers
        if (int rc = check(b->stream ? hipStreamSynchronize(b->stream) : hipDeviceSynchronize(), "sync")) return rc;
        batch_drop_graph(b);
        const int cap = std::max({n, 2 * b->cap, 16});
        batch_free(b);
        if (int rc = check(hipMalloc(&b->table_dev, sizeof(FusedArgs) * cap), "hipMalloc")) return rc;
        if (int rc = check(hipMalloc(&b->meta_dev, sizeof(int) * (cap + 2)), "hipMalloc")) return rc;
        for (int i = 0; i < 2; ++i) {
            if (int rc = check(hipHostMalloc(&b->table_host[i], sizeof(FusedArgs) * cap, 0), "hipHostMalloc"))
                return rc;
            if (int rc = check(hipHostMalloc(&b->meta_host[i], sizeof(int) * (cap + 2), 0), "hipHostMalloc"))
                return rc;
        }
        b->cap = cap;
    }
    // staging buffer i may still feed an upload queued two sets ago
    const int i = b->stage;
    if (b->pending[i])
        if (int rc = check(hipEventSynchronize(b->copied[i]), "staging wait")) return rc;
    FusedArgs* tab = b->table_host[i];
    int* meta = b->meta_host[i];
    int total = 0;
    int64_t cells = 0;
    meta[0] = n;
    meta[1] = 0;
    for (int q = 0; q < n; ++q) {
        FusedArgs a = ctxs[q]->a;
        cells += static_cast<int64_t>(a.nx) * a.ny * a.nz;
        fused_grid(a, b->rows, ctxs[q]->planes, 1);
        a.koff = 0;
        a.step_limit = step_limits[q];
        tab[q] = a;
        total += a.nwaves;
        meta[q + 2] = total;
    }
    if (n > 0)
        if (int rc = check(hipMemcpyAsync(b->table_dev, tab, sizeof(FusedArgs) * n, hipMemcpyHostToDevice, b->stream),
                           "table upload"))
            return rc;
    if (int rc = check(hipMemcpyAsync(b->meta_dev, meta, sizeof(int) * (n + 2), hipMemcpyHostToDevice, b->stream),
                       "offsets upload"))
        return rc;
    if (int rc = check(hipEventRecord(b->copied[i], b->stream), "event record")) return rc;
    b->pending[i] = true;
    b->stage ^= 1;
    b->n = n;
    b->total_waves = total;
    b->nts = cells > kCacheResidentCells;
    return 0;
}

// n_steps launches (step koff = 0 .. n_steps - 1 of the chunk, each counted
// from the solves' counters), then one launch that moves the counters.
static int batch_enqueue(fdtd_batch* b, hipStream_t st, int n_steps, int cap_waves) {
    for (int s = 0; s < n_steps; ++s)
        if (int rc = check(launch_batch(b->table_dev, b->meta_dev + 1, b->meta_dev, cap_waves, b->rows, b->nts, s, st),
                           "batch step launch"))
            return rc;
    return check(launch_batch_advance(b->table_dev, b->meta_dev, b->cap, n_steps, st), "batch advance launch");
}

int fdtd_batch_step_n(fdtd_batch* b, int n_steps, int use_graph) {
    if (!b) return fail(-1, "null batch");
    if (n_steps < 0) return fail(-2, "negative step count");
    if (b->n == 0 || n_steps == 0) return 0;
    DeviceGuard g(b->device);
    if (!use_graph) return batch_enqueue(b, b->stream, n_steps, b->total_waves);
    if (!b->graph || b->graph_steps != n_steps || b->graph_waves < b->total_waves || b->graph_nts != b->nts) {
        // capture with head room so that later tables of up to this many waves
        // replay it; a recapture for another reason keeps the capacity
        int cap_waves = b->graph_waves;
        if (cap_waves < b->total_waves) cap_waves = std::max(b->total_waves + b->total_waves / 2, b->graph_waves * 2);
        batch_drop_graph(b);
        hipStream_t cap = nullptr;
        if (int rc = check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate")) return rc;
        hipGraph_t gr = nullptr;
        if (int rc = check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture")) {
            (void)hipStreamDestroy(cap);
            return rc;
        }
        int rc = batch_enqueue(b, cap, n_steps, cap_waves);
        hipError_t e = hipStreamEndCapture(cap, &gr);
        (void)hipStreamDestroy(cap);
        if (rc) { if (gr) (void)hipGraphDestroy(gr); return rc; }
        if (int r2 = check(e, "end capture")) return r2;
        e = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (int r3 = check(e, "graph instantiate")) return r3;
        b->graph_steps = n_steps;
        b->graph_waves = cap_waves;
        b->graph_nts = b->nts;
    }
    return check(hipGraphLaunch(b->graph, b->stream), "graph launch");
}

int fdtd_batch_info(fdtd_batch* b, int* n, int* total_waves) {
    if (!b) return fail(-1, "null batch");
    if (n) *n = b->n;
    if (total_waves) *total_waves = b->total_waves;
    return 0;
}

// ---- per-phase ADE entry points -------------------------------------------------
static int ade_poles_ok(const char* who, const fdtd_ade_pole* poles, int n, int nx, int ny, int nz) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(-2, "%s: bad dimensions", who);
    if (static_cast<int64_t>(nx) * ny * nz >= INT32_MAX) return fail(-2, "%s: grid too large", who);
    if (n < 0 || (n > 0 && !poles)) return fail(-2, "%s: bad pole list", who);
    for (int q = 0; q < n; ++q) {
        const fdtd_ade_pole& pl = poles[q];
        if (pl.material_id < 0 || pl.material_id > 255 || (pl.target != 0 && pl.target != 1) ||
            (pl.lorentz != 0 && pl.lorentz != 1) || pl.field_index < 0)
            return fail(-2, "%s: pole %d is malformed", who, q);
    }
    return 0;
}

// Poles of one list (lorentz) and target, in the order given: one pass each.
static int ade_update(const char* who, float* J, float* Jp, const float* src, const uint8_t* mat,
                      const fdtd_ade_pole* poles, int n, int lorentz, int target, int nx, int ny, int nz,
                      void* stream) {
    if (int rc = ade_poles_ok(who, poles, n, nx, ny, nz)) return rc;
    if (!J || !src || !mat || (lorentz && !Jp)) return fail(-1, "%s: null device pointer", who);
    const int64_t cells = static_cast<int64_t>(nx) * ny * nz;
    for (int q = 0; q < n; ++q) {
        const fdtd_ade_pole& pl = poles[q];
        if (pl.lorentz != lorentz || pl.target != target) continue;
        const int64_t off = pl.field_index * cells;
        if (int rc = check(phase_ade_update(J + off, lorentz ? Jp + off : nullptr, src, mat, pl.material_id, lorentz,
                                            pl.a, pl.b, pl.d, cells, static_cast<hipStream_t>(stream)),
                           who))
            return rc;
    }
    return 0;
}

int fdtd_update_ade_density_debye(float* J_debye, const float* p, const uint8_t* material, const fdtd_ade_pole* poles,
                                  int n_poles, int nx, int ny, int nz, void* stream) {
    return ade_update("update_ade_density_debye", J_debye, nullptr, p, material, poles, n_poles, 0, 0, nx, ny, nz,
                      stream);
}

int fdtd_update_ade_modulus_debye(float* J_debye, const float* div, const uint8_t* material, const fdtd_ade_pole* poles,
                                  int n_poles, int nx, int ny, int nz, void* stream) {
    return ade_update("update_ade_modulus_debye", J_debye, nullptr, div, material, poles, n_poles, 0, 1, nx, ny, nz,
                      stream);
}

int fdtd_update_ade_density_lorentz(float* J_lorentz, float* J_lorentz_prev, const float* p, const uint8_t* material,
                                    const fdtd_ade_pole* poles, int n_poles, int nx, int ny, int nz, void* stream) {
    return ade_update("update_ade_density_lorentz", J_lorentz, J_lorentz_prev, p, material, poles, n_poles, 1, 0, nx,
                      ny, nz, stream);
}

int fdtd_update_ade_modulus_lorentz(float* J_lorentz, float* J_lorentz_prev, const float* div, const uint8_t* material,
                                    const fdtd_ade_pole* poles, int n_poles, int nx, int ny, int nz, void* stream) {
    return ade_update("update_ade_modulus_lorentz", J_lorentz, J_lorentz_prev, div, material, poles, n_poles, 1, 1, nx,
                      ny, nz, stream);
}

int fdtd_apply_ade_velocity_correction(float* vx, float* vy, float* vz, const float* J_debye, const float* J_lorentz,
                                       const uint8_t* material, const fdtd_ade_pole* poles, int n_poles,
                                       const float* rho_inf, float dt, float inv_dx, int nx, int ny, int nz,
                                       void* stream) {
    const char* who = "apply_ade_velocity_correction";
    if (int rc = ade_poles_ok(who, poles, n_poles, nx, ny, nz)) return rc;
    if (!vx || !vy || !vz || !material || (n_poles > 0 && !rho_inf)) return fail(-1, "%s: null pointer", who);
    const int64_t cells = static_cast<int64_t>(nx) * ny * nz;
    for (int lorentz = 0; lorentz < 2; ++lorentz)
        for (int q = 0; q < n_poles; ++q) {
            const fdtd_ade_pole& pl = poles[q];
            if (pl.lorentz != lorentz || pl.target != 0) continue;
            const float* J = lorentz ? J_lorentz : J_debye;
            if (!J) return fail(-1, "%s: null J array", who);
            // the reference forms the coefficient in float: -dt / rho_inf * inv_dx
            const float coeff = -dt / rho_inf[pl.material_id] * inv_dx;
            if (int rc = check(phase_ade_vcorr(vx, vy, vz, J + pl.field_index * cells, material, pl.material_id, coeff,
                                               nx, ny, nz, static_cast<hipStream_t>(stream)),
                               who))
                return rc;
        }
    return 0;
}

int fdtd_apply_ade_pressure_correction(float* p, const float* J_debye, const float* J_lorentz, const uint8_t* material,
                                       const fdtd_ade_pole* poles, int n_poles, const float* K_inf, float dt, int nx,
                                       int ny, int nz, void* stream) {
    const char* who = "apply_ade_pressure_correction";
    if (int rc = ade_poles_ok(who, poles, n_poles, nx, ny, nz)) return rc;
    if (!p || !material || (n_poles > 0 && !K_inf)) return fail(-1, "%s: null pointer", who);
    const int64_t cells = static_cast<int64_t>(nx) * ny * nz;
    for (int lorentz = 0; lorentz < 2; ++lorentz)
        for (int q = 0; q < n_poles; ++q) {
            const fdtd_ade_pole& pl = poles[q];
            if (pl.lorentz != lorentz || pl.target != 1) continue;
            const float* J = lorentz ? J_lorentz : J_debye;
            if (!J) return fail(-1, "%s: null J array", who);
            const float coeff = -K_inf[pl.material_id] * dt;
            if (int rc = check(phase_ade_pcorr(p, J + pl.field_index * cells, material, pl.material_id, coeff, cells,
                                               static_cast<hipStream_t>(stream)),
                               who))
                return rc;
        }
    return 0;
}

}  // extern "C"
