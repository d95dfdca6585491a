/*
 * fdtd_mi355x.h — C ABI of the MI355X (gfx950) backend for the 3D staggered
 * pressure-velocity FDTD step of strata-fdtd.
 *
 * All field, table and trace pointers are DEVICE pointers (the Python host
 * passes torch.Tensor.data_ptr() of tensors it owns). Every entry point returns
 * 0 on success and a negative code on failure; fdtd_last_error() returns the
 * message of the last failure on the calling thread. Step entry points only
 * enqueue work on the context's stream: they never allocate, never
 * synchronise and are capturable into a hipGraph.
 *
 * Layout: C order (nx, ny, nz), k fastest, flat index (i*ny + j)*nz + k, as
 * the reference's _kernels/include/fdtd_types.hpp:33-42. Staggering: vx[i]
 * sits on the face between cells i and i+1 (core/solver.py:2009-2013).
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef FDTD_MI355X_H
#define FDTD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cell-flag byte built on the device from the geometry mask (bit layout). */
#define FDTD_FLAG_AIR    0x01u  /* cell is air (geometry True)                  */
#define FDTD_FLAG_OPEN_X 0x02u  /* face (i|i+1) below the last plane, both air  */
#define FDTD_FLAG_OPEN_Y 0x04u
#define FDTD_FLAG_OPEN_Z 0x08u
#define FDTD_FLAG_ADE    0x10u  /* cell carries a material with poles          */

/* Limits. Sources: the point and x/y-plane sources the sweep adds itself per
 * context; the host routes any further source to the per-cell tables
 * (fdtd_ctx_set_membranes), so a solve's source list is unbounded. Probes:
 * the first 64 are recorded inside the sweep, the rest by a small kernel after
 * the step. Poles: read from a device table. */
#define FDTD_MAX_SOURCES 8
#define FDTD_MAX_PROBES  (1 << 20)
#define FDTD_MAX_MICS 4096
#define FDTD_MAX_POLES   1024

typedef struct fdtd_ctx fdtd_ctx;

/* ---- errors / info ------------------------------------------------------ */
/* Replaces: RuntimeError messages of _kernels/kernels.cpp:22-65. */
const char* fdtd_last_error(void);
/* Replaces: _kernels.__version__ / has_openmp / get_num_threads (kernels.cpp:74-102). */
const char* fdtd_version(void);
int fdtd_device_count(int* out);

/* ---- context (one solve, one stream) ------------------------------------ */
/* Replaces: FDTDSolver's _use_native state (core/solver.py:1536-1574). stream
 * is a hipStream_t (NULL = legacy default stream). */
int fdtd_ctx_create(int device, void* stream, fdtd_ctx** out);
int fdtd_ctx_destroy(fdtd_ctx* ctx);

/* Grid and coefficients. Replaces: FDTDSolver.__init__ coefficient setup
 * (core/solver.py:1582-1602) and create_nonuniform_grid_data (kernels.cpp:443-475).
 * For a uniform grid pass nonuniform=0 and NULL tables; cv/cp are
 * coeff_v/coeff_p. For a nonuniform grid cv/cp are coeff_v_base/coeff_p_base and
 * inv_face[a] (n_a - 1 floats) / inv_cell[a] (n_a floats) are device tables. */
int fdtd_ctx_set_grid(fdtd_ctx* ctx, int nx, int ny, int nz, int nonuniform, float cv, float cp,
                      const float* inv_face_x, const float* inv_face_y, const float* inv_face_z,
                      const float* inv_cell_x, const float* inv_cell_y, const float* inv_cell_z);

/* Ping-pong field buffers: set[0] = {p, vx, vy, vz}, set[1] likewise. Step s
 * reads set[s & 1] and writes set[(s & 1) ^ 1]. Replaces the in-place numpy
 * arrays solver.p/vx/vy/vz (core/solver.py:1604-1608). */
int fdtd_ctx_set_fields(fdtd_ctx* ctx, float* p0, float* vx0, float* vy0, float* vz0, float* p1,
                        float* vx1, float* vy1, float* vz1);

/* Cell flags (1 byte/cell, see FDTD_FLAG_*). Replaces precompute_boundary_cells
 * + the bool geometry argument (boundaries.cpp:13-64, fdtd_step.cpp:84-91). */
int fdtd_ctx_set_flags(fdtd_ctx* ctx, const uint8_t* flags);

/* PML decay tables exp(-sigma dt) per axis (device, n_a floats; tables of
 * 1.0f for an axis without PML). Replaces initialize_pml / PMLData
 * (pml.cpp:13-45). */
int fdtd_ctx_set_pml(fdtd_ctx* ctx, const float* decay_x, const float* decay_y,
                     const float* decay_z);

/* Sources: kind[q] 0 = point (a[q] = flat cell index), 1 = plane (a[q] = axis,
 * b[q] = index). wave: device float64 table [n][wave_len], value at step s in
 * column s. Host arrays for kind/a/b. Replaces _inject_sources
 * (core/solver.py:2386-2433). */
int fdtd_ctx_set_sources(fdtd_ctx* ctx, int n, const int32_t* kind, const int64_t* a,
                         const int64_t* b, const double* wave, int64_t wave_len);

/* Probes: host flat indices; trace: device float32 [trace_len][n]; step s
 * writes row s. Replaces _record_probes (core/solver.py:2435-2439). */
int fdtd_ctx_set_probes(fdtd_ctx* ctx, int n, const int64_t* flat_idx, float* trace,
                        int64_t trace_len);

/* Microphones: n (<= FDTD_MAX_MICS) trilinear gathers recorded after every
 * step into out [out_len][n][4] (p, vx, vy, vz at the microphone, float64).
 * idx: host int32 [n][4][8] flat cell indices; w: host float64 [n][4][8]
 * weights (corner order and weights of Microphone._compute_trilinear_weights,
 * core/solver.py:962-1003; velocity rows as _interpolate_velocity :1060-1113);
 * f64: host int32 [n][4], 1 = sum in float64 (the reference's weights are
 * numpy float64), 0 = weights rounded to float32 and summed in float32 (the
 * reference's weights are Python floats).
 * Replaces _record_microphones (core/solver.py:2441-2518) and the native
 * precompute_microphone_data / record_microphones_batch
 * (_kernels/src/microphones.cpp:16-116). The tables are copied to the device. */
int fdtd_ctx_set_microphones(fdtd_ctx* ctx, int n, const int32_t* idx, const double* w, const int32_t* f64,
                             double* out, int64_t out_len);

/* Per-cell source tables (membranes, and any point or plane source a caller
 * routes here with weight 1.0), applied after each step's sweep, ADE fix-up
 * and the sweep's own point/plane sources: for every listed cell (flat index
 * below 2^31, cells increasing), entries entry_start[e] .. entry_start[e+1]-1
 * in source order
 * do field[c] = float(double(field[c]) + wave[src][s] * weight), field 0 p,
 * 1 vx, 2 vy, 3 vz; then the probes probe_idx[probe_start[e] ..
 * probe_start[e+1]-1] record p at the cell again. Host arrays except wave
 * (device float64 [n_src][wave_len], value at step s in column s); the
 * tables are copied to the device after the steps already enqueued on the
 * context stream finish. n_cells = 0 removes them. Replaces the membrane
 * branch of _inject_sources (core/solver.py:2389-2412), and the point/plane
 * branches for the sources a caller lists here. */
int fdtd_ctx_set_membranes(fdtd_ctx* ctx, int n_cells, const int64_t* cell, const int32_t* entry_start,
                           int n_entries, const int32_t* entry_src, const int32_t* entry_field,
                           const double* entry_weight, const int32_t* probe_start, int n_probe_refs,
                           const int32_t* probe_idx, int n_src, const double* wave, int64_t wave_len);

/* Boundary planes updated after each step's pressure update (and PML damping),
 * before its sources and probes, in list order. Replaces apply_pressure of
 * ABCFirstOrder (first-order Mur, boundaries/_boundaries.py:475-511) and of
 * RadiationImpedance (:707-776); both exist only in python in the reference.
 *   kind 0, Mur, both faces of `axis`:
 *     p[0] = prev0 + coeff (p[1] - p[0]), prev0 = p[1];
 *     p[N-1] = prev1 + coeff (p[N-2] - p[N-1]), prev1 = p[N-2]
 *   kind 1, radiation impedance, face `side` (0 low, 1 high), inner plane p_in:
 *     abc = prev + coeff (p_in - p_b); p_b = refl p_in + omr abc; prev = p_in
 * with numpy's types: float32 differences, float64 products and sums with the
 * float64 coeff and omr (= 1 - R), refl rounded to float32 for refl * p_in,
 * one rounding to float32 per stored value. */
typedef struct {
    int32_t kind;        /* 0 Mur ABC, 1 radiation impedance */
    int32_t axis;        /* 0 x, 1 y, 2 z */
    int32_t side;        /* kind 1: 0 low face, 1 high face */
    int32_t pad;
    double coeff;        /* Mur coefficient (c dt - dx) / (c dt + dx) */
    double refl;         /* kind 1: reflection coefficient R */
    double omr;          /* kind 1: 1 - R */
    int64_t state_off;   /* first element of this entry's previous-step planes in `state` */
} fdtd_boundary_plane;
#define FDTD_MAX_BOUNDARY_PLANES 16
/* planes: host array of n entries (n <= FDTD_MAX_BOUNDARY_PLANES; 0 removes
 * them). state: device float32 [state_len], zeroed by the caller, holding each
 * entry's previous-step planes (a Mur entry two cross-sections of its axis,
 * a radiation entry one) from its state_off on; the library keeps using it.
 * Every axis a plane acts on needs at least 2 cells. While planes are set,
 * every probe is recorded after the step's last pass and the step kernels add
 * no sources themselves: the caller lists every source in the per-cell tables
 * (fdtd_ctx_set_membranes), which run after the planes, as the reference adds
 * its sources after its boundaries (core/solver.py:2044-2049). */
int fdtd_ctx_set_boundary_planes(fdtd_ctx* ctx, int n, const fdtd_boundary_plane* planes, float* state,
                                 int64_t state_len);

/* Device int64 holding the index of the next step (read by every step kernel,
 * advanced by fdtd_step_n) and a device uint32 health word (bit0: a probe
 * sample was not finite). */
int fdtd_ctx_set_counters(fdtd_ctx* ctx, int64_t* step_counter, uint32_t* health);

/* ADE (frequency-dependent materials). Replaces ADEMaterialData + the
 * update_ade_* / apply_ade_* kernels (ade.cpp:25-473). Poles are host arrays:
 * mat_id, target (0 density, 1 modulus), lorentz (0/1), coefficients a,b,d and
 * corr (density: -dt/rho_inf * inv_dx, with inv_dx = 1/min spacing on every
 * grid as the native correction is called, core/solver.py:3295-3305; modulus:
 * -K_inf dt). The
 * auxiliary fields live on a dense box [lo, hi) (the material bounding box):
 * J and Jprev are device arrays [n_pole][box cells]. material: device uint8
 * per grid cell; the call waits for the device and reads the box's ids once to
 * list the cells whose material has poles (the ADE kernels visit only those),
 * so call it again after changing the ids. */
int fdtd_ctx_set_ade(fdtd_ctx* ctx, int n_pole, const int32_t* mat_id, const int32_t* target,
                     const int32_t* lorentz, const float* a, const float* b, const float* d,
                     const float* corr, const uint8_t* material, const int32_t* box_lo,
                     const int32_t* box_hi, float* J, float* Jprev, float inv_dx);

/* Launch shape for the fused kernel: rows per wave (2, 4, 8 or 16), planes
 * per wave (>= 1) and stack, the number of waves (1, 2, 4 or 8) of one
 * workgroup that own consecutive j-strips and share their halo rows through
 * LDS (1: independent waves). 0 picks the built-in default (4, 8, 1). No
 * reference counterpart (the reference's OpenMP loops have no launch shape). */
int fdtd_ctx_set_tiling(fdtd_ctx* ctx, int rows, int planes, int stack);

/* Enqueue n fused time steps. Replaces FDTDSolver.step's native dispatch
 * (core/solver.py:2028-2077: _step_native + boundary.apply_* + sources +
 * probes). Step s (read from the device counter) reads buffer set s&1. */
int fdtd_step_n(fdtd_ctx* ctx, int n_steps);

/* Capture n fused steps (plus the counter advance) into a hipGraph owned by
 * the context and replay it `replays` times. */
int fdtd_graph_run(fdtd_ctx* ctx, int n_steps, int replays);

/* Average device time of the fused step kernel over the last timed window,
 * measured with HIP events on the context stream (milliseconds). */
int fdtd_time_steps(fdtd_ctx* ctx, int n_steps, int use_graph, float* ms_total);

/* Enqueue n_steps (1..4096) eager steps with HIP events recorded around each
 * fused sweep launch, wait, and return the summed sweep time and the whole
 * window (milliseconds). Timing instrumentation; no reference counterpart. */
int fdtd_time_sweeps(fdtd_ctx* ctx, int n_steps, float* ms_sweep, float* ms_total);

/* Practical bandwidth ceiling of the fused sweep's access mix: a streaming
 * kernel that reads p, vx, vy, vz and the flag byte of every cell from the
 * current buffer set and writes the four fields to the other set (the next
 * step's output, so the solve's state is unchanged), with the sweep's store
 * kind. One warm-up pass, then `reps` (1..1000) timed passes; returns the
 * average milliseconds per pass. Field buffers must be 16-byte aligned.
 * Measurement instrumentation; no reference counterpart. */
int fdtd_stream_ceiling(fdtd_ctx* ctx, int reps, float* ms_per_pass);

/* The same streaming work in the fused sweep's own order: one lane per k
 * cell, the context's rows x planes strips per wave and the sweep's block
 * remap, with no stencil and no halo; the sweep's cost beyond it is its
 * stencil and halo traffic. Grids below 2^30 cells, rows 2, 4 or 8. Timing as
 * fdtd_stream_ceiling. Measurement instrumentation; no reference counterpart. */
int fdtd_stream_ceiling_march(fdtd_ctx* ctx, int reps, float* ms_per_pass);

/* ---- observation of the state (between steps, on the context's stream) -- */
/* Both read the buffer set the next step reads (set counter & 1 of the device
 * step counter), so they order after the steps already enqueued and need no
 * host sync; outputs are device pointers. */

/* Energy sums of the current state: out[0] = sum over air cells of p^2,
 * out[1] = sum over air cells of vx^2 + vy^2 + vz^2, in float64 (every
 * float32 square is exact in float64), formed in a fixed order so that a call
 * repeats to the bit. The host forms 1/2 out[0]/(rho c^2) dV + 1/2 rho out[1]
 * dV. Field buffers 16-byte and flags 4-byte aligned. Replaces
 * FDTDSolver.compute_energy (core/solver.py:2689-2706). */
int fdtd_ctx_energy(fdtd_ctx* ctx, double* out);

/* Cell-centred velocities (three device float32 arrays of nx*ny*nz):
 * out[c] = 0.5 * (v[c - e_axis] + v[c]), and 0.5 * v[c] on the first plane
 * of the axis. Replaces FDTDSolver._interpolate_v{x,y,z}_to_centers
 * (core/solver.py:2647-2687), used by velocity snapshots. */
int fdtd_ctx_centred_velocity(fdtd_ctx* ctx, float* vx, float* vy, float* vz);

/* ---- batched serving (no counterpart in the reference) ------------------- */
/* A batch advances many independent solves (uniform grids, no ADE) by one step
 * per launch: waves are dealt out over the solves' strips, each solve keeps its
 * own buffers, tables and counter, and a solve whose counter reached its
 * limit is left untouched. Replaces running BatchedGPUFDTDSolver-style band
 * batches (core/solver_gpu.py:978-1351) for a design sweep. */
typedef struct fdtd_batch fdtd_batch;
int fdtd_batch_create(int device, void* stream, int rows, fdtd_batch** out);
int fdtd_batch_destroy(fdtd_batch* b);
/* (Re)build the device table from n ready contexts; step_limits[q] is the
 * total step count of solve q. Set-up call (may allocate, synchronises). */
int fdtd_batch_set(fdtd_batch* b, fdtd_ctx* const* ctxs, const int64_t* step_limits, int n);
/* Enqueue n batched steps (use_graph: capture once per size, then replay):
 * n launches, step k of each solve's chunk counted from its counter, then one
 * small launch that moves every counter by n, up to its limit. */
int fdtd_batch_step_n(fdtd_batch* b, int n_steps, int use_graph);
int fdtd_batch_info(fdtd_batch* b, int* n, int* total_waves);

/* ---- per-voxel density / modulus and the adjoint (no reference counterpart) */
/* The reference has scalar rho and c (core/solver.py:1584-1592). These two
 * steps add per-voxel r = rho/rho0 and kappa = K/K0 on a design box [lo, hi)
 * (float32 box arrays; NULL = 1 everywhere) and the time-reversed adjoint.
 * dtype 4 = float32 fields, 8 = float64 fields. One step per call, on `stream`.
 * fdtd_param_step: (p, v) -> (po, vo), probe samples into trace_row[n_probe],
 * source value w added (float64) at src_cell (-1: none).
 * fdtd_adjoint_step: (P, V) = adjoint at step n+1 -> (Po, Vo) at step n, with
 * the forward state n in (p, v), dL/dy_n in g_row[n_probe]; accumulates
 * dL/dr and dL/dkappa on the box with Kahan compensation (sum, comp pairs). */
int fdtd_param_step(int dtype, const void* p, const void* vx, const void* vy, const void* vz, void* po,
                    void* vxo, void* vyo, void* vzo, int nx, int ny, int nz, const uint8_t* flags,
                    const float* dx, const float* dy, const float* dz, const int32_t* box_lo,
                    const int32_t* box_hi, const float* r, const float* kappa, double cv0, double cp0,
                    int64_t src_cell, double w, int n_probe, const int64_t* probe_cell, void* trace_row,
                    void* stream);
int fdtd_adjoint_step(int dtype, const void* P, const void* Vx, const void* Vy, const void* Vz, void* Po,
                      void* Vxo, void* Vyo, void* Vzo, const void* p, const void* vx, const void* vy,
                      const void* vz, int nx, int ny, int nz, const uint8_t* flags, const float* dx,
                      const float* dy, const float* dz, const int32_t* box_lo, const int32_t* box_hi,
                      const float* r, const float* kappa, double cv0, double cp0, int n_probe,
                      const int64_t* probe_cell, const void* g_row, void* gr_sum, void* gr_comp,
                      void* gk_sum, void* gk_comp, void* stream);
/* Many steps per call (the loops run in C, not per step from Python).
 * fdtd_param_run: states = host array of 4 * n_states device pointers
 * {p, vx, vy, vz}; pingpong: step m reads states[m & 1], writes
 * states[(m + 1) & 1] (n_states = 2), else reads states[m], writes
 * states[m + 1] (n_states = n_steps + 1); wave: host float64 waveform by
 * absolute step; trace: device [steps][n_probe], row step0 + m.
 * fdtd_adjoint_run: lam = host array of 8 device pointers (two {P, Vx, Vy, Vz}
 * buffers, lambda_{n_hi} in buffer 0; lambda_{n_lo} ends in buffer
 * (n_hi - n_lo) & 1); states = 4 * (n_hi - n_lo) pointers, forward state of
 * step n_lo + m at entry m; g = device [steps][n_probe]. */
int fdtd_param_run(int dtype, void* const* states, int n_states, int64_t step0, int n_steps, int pingpong,
                   int nx, int ny, int nz, const uint8_t* flags, const float* dx, const float* dy,
                   const float* dz, const int32_t* box_lo, const int32_t* box_hi, const float* r,
                   const float* kappa, double cv0, double cp0, int64_t src_cell, const double* wave,
                   int n_probe, const int64_t* probe_cell, void* trace, void* stream);
int fdtd_adjoint_run(int dtype, void* const* lam, void* const* states, int64_t n_lo, int64_t n_hi, int nx,
                     int ny, int nz, const uint8_t* flags, const float* dx, const float* dy, const float* dz,
                     const int32_t* box_lo, const int32_t* box_hi, const float* r, const float* kappa,
                     double cv0, double cp0, int n_probe, const int64_t* probe_cell, const void* g,
                     void* gr_sum, void* gr_comp, void* gk_sum, void* gk_comp, void* stream);

/* Fast float32 design solve (2.5D marching sweeps, same per-cell arithmetic
 * as fdtd_param_step / fdtd_adjoint_step). fields / lam: host arrays of 8
 * device pointers (two {p, vx, vy, vz} sets; step s reads set s & 1, the
 * adjoint step n reads set (n + 1) & 1). dec: host array of 3 device decay
 * tables. Grids below 2^30 cells; the design box must keep one cell to every
 * grid face. rk: device float32 [hx][hy][hz][2], (rho/rho0, K/K0) per box
 * voxel, or NULL (air). src_ijk: host int32[3] or NULL; wave: device float64
 * [steps_len]; probe_ijk: device int32 [n_probe][3]; trace and g: device
 * float32 [steps_len][n_probe], indexed by absolute step (calls that would
 * touch a row at or past steps_len return -4).
 * History (device float32): p on the box grown by one cell
 * ((hx+2)(hy+2)(hz+2) per step) and div on the box (hx hy hz per step) for
 * steps [hist0, hist0 + hist_steps); fdtd_design_history_bytes gives the size.
 * The forward records it when hist_p is not NULL; the adjoint (steps
 * n_hi - 1 down to n_lo) reads it and Kahan-accumulates into acc, device
 * float32 [hx][hy][hz][4]: (sum, compensation) of dL/d(rho/rho0), then of
 * dL/d(K/K0); the gradient is sum - compensation. face_e: device float32
 * scratch of fdtd_design_face_count entries (every box face's gradient
 * factor, formed once per call); pb_box: device float32 scratch of
 * 2 hx hy hz entries (each step's adjoint pressure on the box, kept for the
 * gradient pass, which takes two steps at a time). */
int64_t fdtd_design_face_count(const int32_t* box_lo, const int32_t* box_hi);
int64_t fdtd_design_history_bytes(const int32_t* box_lo, const int32_t* box_hi, int64_t steps);
int fdtd_design_forward(float* const* fields, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* rk,
                        const int32_t* src_ijk, const double* wave, int n_probe, const int32_t* probe_ijk,
                        float* trace, int64_t steps_len, int64_t step0, int n_steps, float* hist_p, float* hist_div,
                        int64_t hist0, int64_t hist_steps, int planes, void* stream);
int fdtd_design_adjoint(float* const* lam, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* rk,
                        int n_probe, const int32_t* probe_ijk, const float* g, int64_t steps_len, int64_t n_lo,
                        int64_t n_hi, const float* hist_p, const float* hist_div, int64_t hist0, int64_t hist_steps,
                        float* acc, float* face_e, float* pb_box, int planes, void* stream);

/* ---- per-phase entry points (one kernel per reference function) --------- */
/* All in place, on `stream`, float32, uniform unless noted. */
/* Replaces fdtd::update_velocity (fdtd_step.cpp:15-82). */
int fdtd_update_velocity(const float* p, float* vx, float* vy, float* vz, int nx, int ny, int nz,
                         float cv, void* stream);
/* Replaces fdtd::update_velocity_nonuniform (fdtd_step.cpp:235-307). */
int fdtd_update_velocity_nonuniform(const float* p, float* vx, float* vy, float* vz, int nx,
                                    int ny, int nz, float cv_base, const float* ifx,
                                    const float* ify, const float* ifz, void* stream);
/* Replaces fdtd::apply_rigid_boundaries (boundaries.cpp:66-89), driven by flags. */
int fdtd_apply_rigid(float* vx, float* vy, float* vz, const uint8_t* flags, int nx, int ny,
                     int nz, void* stream);
/* Replaces fdtd::update_pressure (fdtd_step.cpp:84-212). */
int fdtd_update_pressure(float* p, const float* vx, const float* vy, const float* vz,
                         const uint8_t* flags, int nx, int ny, int nz, float cp, void* stream);
/* Replaces fdtd::update_pressure_nonuniform (fdtd_step.cpp:309-439). */
int fdtd_update_pressure_nonuniform(float* p, const float* vx, const float* vy, const float* vz,
                                    const uint8_t* flags, int nx, int ny, int nz, float cp_base,
                                    const float* icx, const float* icy, const float* icz,
                                    void* stream);
/* Replaces fdtd::apply_pml_velocity / apply_pml_pressure (pml.cpp:47-149).
 * NULL table = axis without PML. */
int fdtd_apply_pml_velocity(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                            const float* dz, int nx, int ny, int nz, void* stream);
int fdtd_apply_pml_pressure(float* p, const float* dx, const float* dy, const float* dz, int nx,
                            int ny, int nz, void* stream);
/* Replaces fdtd::compute_divergence(_nonuniform) (ade.cpp:479-692); icx NULL
 * = uniform with inv_dx. */
int fdtd_compute_divergence(float* div, const float* vx, const float* vy, const float* vz, int nx,
                            int ny, int nz, float inv_dx, const float* icx, const float* icy,
                            const float* icz, void* stream);
/* ---- per-phase ADE entry points ------------------------------------------
 * One pole of ADEMaterialData (_kernels/include/ade.hpp:44-111): lorentz 0
 * puts it in the Debye list, 1 in the Lorentz list; field_index is its row in
 * the J array of that list (rows of nx*ny*nz floats, J_prev alike); a, b, d
 * are the Debye (alpha, beta, -) or Lorentz (a, b, d) coefficients. poles is a
 * host array; each call visits the poles of its list and target in the order
 * given, as the reference iterates its vectors, one device pass per pole.
 * material: device uint8 ids per cell. rho_inf / K_inf: host float[256] by
 * material id. In place on `stream`, float32. The fused step applies the same
 * updates in its own launches; these exist for phase-by-phase bisection. */
typedef struct fdtd_ade_pole {
    int32_t material_id, target, lorentz, field_index;
    float a, b, d, reserved;
} fdtd_ade_pole;
/* Replaces fdtd::update_ade_density_debye (ade.cpp:25-68; kernels.cpp:670-695). */
int fdtd_update_ade_density_debye(float* J_debye, const float* p, const uint8_t* material,
                                  const fdtd_ade_pole* poles, int n_poles, int nx, int ny, int nz, void* stream);
/* Replaces fdtd::update_ade_modulus_debye (ade.cpp:70-112; kernels.cpp:697-720). */
int fdtd_update_ade_modulus_debye(float* J_debye, const float* div, const uint8_t* material,
                                  const fdtd_ade_pole* poles, int n_poles, int nx, int ny, int nz, void* stream);
/* Replaces fdtd::update_ade_density_lorentz (ade.cpp:118-170; kernels.cpp:722-748). */
int fdtd_update_ade_density_lorentz(float* J_lorentz, float* J_lorentz_prev, const float* p, const uint8_t* material,
                                    const fdtd_ade_pole* poles, int n_poles, int nx, int ny, int nz, void* stream);
/* Replaces fdtd::update_ade_modulus_lorentz (ade.cpp:172-222; kernels.cpp:750-776). */
int fdtd_update_ade_modulus_lorentz(float* J_lorentz, float* J_lorentz_prev, const float* div,
                                    const uint8_t* material, const fdtd_ade_pole* poles, int n_poles, int nx, int ny,
                                    int nz, void* stream);
/* Replaces fdtd::apply_ade_velocity_correction (ade.cpp:228-401; kernels.cpp:778-826):
 * Debye then Lorentz density poles, coeff = -dt / rho_inf[m] * inv_dx in float. */
int fdtd_apply_ade_velocity_correction(float* vx, float* vy, float* vz, const float* J_debye, const float* J_lorentz,
                                       const uint8_t* material, const fdtd_ade_pole* poles, int n_poles,
                                       const float* rho_inf, float dt, float inv_dx, int nx, int ny, int nz,
                                       void* stream);
/* Replaces fdtd::apply_ade_pressure_correction (ade.cpp:403-473; kernels.cpp:828-858):
 * Debye then Lorentz modulus poles, coeff = -K_inf[m] * dt in float. */
int fdtd_apply_ade_pressure_correction(float* p, const float* J_debye, const float* J_lorentz, const uint8_t* material,
                                       const fdtd_ade_pole* poles, int n_poles, const float* K_inf, float dt, int nx,
                                       int ny, int nz, void* stream);

/* Geometry mask (uint8 0/1) + material ids (nullable) -> cell flags.
 * Replaces precompute_boundary_cells (boundaries.cpp:13-64). */
int fdtd_build_flags(const uint8_t* air, const uint8_t* material, uint8_t* flags, int nx, int ny,
                     int nz, void* stream);

/* ---- device voxeliser ------------------------------------------------------
 * Evaluates CSG programs (boxes, spheres, axis-aligned cylinders and cones;
 * union, intersection, difference; translations and scales of subtrees, at
 * most 8 nested, scale factors in [1e-6, 1e6]; geometry/sdf.py:274-722,
 * 1199-1301) in float64 at the cell centres xc/yc/zc (device)
 * and writes air[c] = (sdf <= 0) and, for material programs, ids[c].
 * Replaces SDFPrimitive.voxelize (geometry/sdf.py:99-127) and
 * MaterializedGeometry.voxelize_with_materials
 * (geometry/material_assignment.py:133-168) with bit-identical masks.
 * ops: host int32 [n_ops][2] postfix instructions ({op, arg}); prog_start:
 * host int32 [n_prog + 1]; prog_ids: host int32 [n_prog] (entry 0 unused);
 * params: host float64 [n_params]. The programs are validated, copied into
 * dev_scratch (device, >= fdtd_voxelize_scratch_bytes bytes; the call waits
 * for that copy) and evaluated on `stream`. ids may be null when n_prog == 1. */
int64_t fdtd_voxelize_scratch_bytes(int n_ops, int n_prog, int n_params);
int fdtd_voxelize(const int32_t* ops, int n_ops, const int32_t* prog_start, const int32_t* prog_ids,
                  int n_prog, const double* params, int n_params, void* dev_scratch, int64_t scratch_bytes,
                  const double* xc, const double* yc, const double* zc, int nx, int ny, int nz,
                  uint8_t* air, uint8_t* ids, void* stream);

/* fdtd_voxelize without the dev_scratch workspace, for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * programs are staged in a fixed device array of FDTD_VOXELIZE_STATIC_BYTES
 * bytes inside the library, one copy per GPU, kept until the process ends;
 * programs needing more staging (fdtd_voxelize_scratch_bytes) are rejected
 * with an error. The call allocates nothing. The GPU is the stream's (the
 * current device for a NULL stream). Every call on one GPU shares its array:
 * calls that overlap on one GPU (other host threads, or other streams whose
 * kernels have not finished) race; calls on one stream are ordered. */
#define FDTD_VOXELIZE_STATIC_BYTES (256 * 1024)
int fdtd_voxelize_static(const int32_t* ops, int n_ops, const int32_t* prog_start, const int32_t* prog_ids,
                         int n_prog, const double* params, int n_params, const double* xc, const double* yc,
                         const double* zc, int nx, int ny, int nz, uint8_t* air, uint8_t* ids, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_MI355X_H */
