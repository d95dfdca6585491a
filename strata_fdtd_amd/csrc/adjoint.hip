// Per-voxel density / bulk-modulus forward step and its hand-written
// time-reversed adjoint (gfx950). No counterpart in the reference: the
// reference has scalar rho and c only (core/solver.py:1584-1592). With
// r = rho/rho0 = 1 and kappa = K/K0 = 1 everywhere the forward step below is,
// operation for operation, the reference step (velocity, rigid, pressure, PML,
// source, probe), so it reduces to the plain kernel bit for bit.
//
// Forward model on a uniform grid (face f between cells lo and hi):
//   CV_f = cv0 / ((r_lo + r_hi) * 0.5)          (arithmetic-mean face density)
//   v'_f = open_f ? v_f + CV_f (p_hi - p_lo) : 0 (faces below the last plane;
//          the last face of each axis passes v through)
//   p'_c = air_c ? p_c + (cp0 kappa_c) div_c : 0 div_c = (dvx + dvy) + dvz
//   v''  = v' * decay ; p'' = ((p' dX) dY) dZ ; p''[src] += w ; y = p''[probe]
// Adjoint of one step (lambda_{n+1} = (P, V) -> lambda_n), with
//   pb_c = dX dY dZ (P_c + g_c)     g: dL/dy injected at the probe cells
//   q_c  = air_c cp0 kappa_c pb_c
//   w_f  = open_f (decay_f V_f + q_lo - q_hi)   (last face: decay V + q_lo, no mask)
//   V_n,f = w_f ; P_n,c = air_c pb_c + sum_{f: c = hi} CV_f w_f - sum_{f: c = lo} CV_f w_f
// and the parameter gradients at step n (forward state n):
//   dL/dkappa_c += air_c pb_c cp0 div'_c
//   dL/dr_c     += sum_{faces f of c} w_f (p_hi - p_lo) (-CV_f / (r_lo + r_hi))
// accumulated over time with Kahan compensation.

#include <algorithm>

#include "fdtd_common.hpp"

namespace fdtd_amd {

struct ParamBox {
    int lo[3], hi[3];
    const float* r;      // rho / rho0 on the box (nullptr: 1 everywhere)
    const float* kappa;  // K / K0 on the box
};

__device__ __forceinline__ bool in_box(const ParamBox& b, int i, int j, int k) {
    return i >= b.lo[0] && i < b.hi[0] && j >= b.lo[1] && j < b.hi[1] && k >= b.lo[2] && k < b.hi[2];
}

__device__ __forceinline__ int64_t box_lin(const ParamBox& b, int i, int j, int k) {
    return ((static_cast<int64_t>(i - b.lo[0]) * (b.hi[1] - b.lo[1])) + (j - b.lo[1])) *
               (b.hi[2] - b.lo[2]) + (k - b.lo[2]);
}

template <typename T>
__device__ __forceinline__ T r_at(const ParamBox& b, int i, int j, int k) {
    return (b.r && in_box(b, i, j, k)) ? static_cast<T>(b.r[box_lin(b, i, j, k)]) : T(1);
}

template <typename T>
__device__ __forceinline__ T kappa_at(const ParamBox& b, int i, int j, int k) {
    return (b.kappa && in_box(b, i, j, k)) ? static_cast<T>(b.kappa[box_lin(b, i, j, k)]) : T(1);
}

struct StepGeom {
    int nx, ny, nz;
    const uint8_t* flags;
    const float* dec[3];
    int64_t src_cell;          // -1: none
    int n_probe;
    const int64_t* probe_cell; // device [n_probe]
};

// One forward step from (p, v) into (po, vo); probe samples into trace_row.
template <typename T>
__global__ void k_param_step(const T* __restrict__ p, const T* __restrict__ vx, const T* __restrict__ vy,
                             const T* __restrict__ vz, T* __restrict__ po, T* __restrict__ vxo,
                             T* __restrict__ vyo, T* __restrict__ vzo, StepGeom g, ParamBox pb, T cv0,
                             T cp0, double w, T* __restrict__ trace_row) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz, n = sx * nx;
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int k = static_cast<int>(c % nz);
        const int j = static_cast<int>((c / nz) % ny);
        const int i = static_cast<int>(c / sx);
        const uint32_t f = g.flags[c];
        const T pc = p[c];
        const T rc = r_at<T>(pb, i, j, k);
        // own faces
        T vxn = vx[c], vyn = vy[c], vzn = vz[c];
        if (i + 1 < nx) {
            const T cvf = cv0 / ((rc + r_at<T>(pb, i + 1, j, k)) * T(0.5));
            vxn = (f & FDTD_FLAG_OPEN_X) ? vxn + cvf * (p[c + sx] - pc) : T(0);
        }
        if (j + 1 < ny) {
            const T cvf = cv0 / ((rc + r_at<T>(pb, i, j + 1, k)) * T(0.5));
            vyn = (f & FDTD_FLAG_OPEN_Y) ? vyn + cvf * (p[c + sy] - pc) : T(0);
        }
        if (k + 1 < nz) {
            const T cvf = cv0 / ((rc + r_at<T>(pb, i, j, k + 1)) * T(0.5));
            vzn = (f & FDTD_FLAG_OPEN_Z) ? vzn + cvf * (p[c + 1] - pc) : T(0);
        }
        // faces below
        T vxm = T(0), vym = T(0), vzm = T(0);
        if (i > 0) {
            const T cvf = cv0 / ((r_at<T>(pb, i - 1, j, k) + rc) * T(0.5));
            vxm = (g.flags[c - sx] & FDTD_FLAG_OPEN_X) ? vx[c - sx] + cvf * (pc - p[c - sx]) : T(0);
        }
        if (j > 0) {
            const T cvf = cv0 / ((r_at<T>(pb, i, j - 1, k) + rc) * T(0.5));
            vym = (g.flags[c - sy] & FDTD_FLAG_OPEN_Y) ? vy[c - sy] + cvf * (pc - p[c - sy]) : T(0);
        }
        if (k > 0) {
            const T cvf = cv0 / ((r_at<T>(pb, i, j, k - 1) + rc) * T(0.5));
            vzm = (g.flags[c - 1] & FDTD_FLAG_OPEN_Z) ? vz[c - 1] + cvf * (pc - p[c - 1]) : T(0);
        }
        const T div = ((vxn - vxm) + (vyn - vym)) + (vzn - vzm);
        T pn = (f & FDTD_FLAG_AIR) ? pc + (cp0 * kappa_at<T>(pb, i, j, k)) * div : T(0);
        const T dx = g.dec[0][i], dy = g.dec[1][j], dz = g.dec[2][k];
        pn = ((pn * dx) * dy) * dz;
        if (c == g.src_cell && (f & FDTD_FLAG_AIR)) pn = static_cast<T>(static_cast<double>(pn) + w);
        po[c] = pn;
        vxo[c] = vxn * dx;
        vyo[c] = vyn * dy;
        vzo[c] = vzn * dz;
        for (int q = 0; q < g.n_probe; ++q)
            if (g.probe_cell[q] == c) trace_row[q] = pn;
    }
}

template <typename T>
__device__ __forceinline__ void kahan_add(T* sum, T* comp, int64_t idx, T x) {
    const T y = x - comp[idx];
    const T t = sum[idx] + y;
    comp[idx] = (t - sum[idx]) - y;
    sum[idx] = t;
}

// One adjoint step: (P, V) = lambda_{n+1} -> (Po, Vo) = lambda_n, with the
// forward state n in (p, vx, vy, vz) and dL/dy_n in g_row (one per probe).
template <typename T>
__global__ void k_adjoint_step(const T* __restrict__ P, const T* __restrict__ Vx, const T* __restrict__ Vy,
                               const T* __restrict__ Vz, T* __restrict__ Po, T* __restrict__ Vxo,
                               T* __restrict__ Vyo, T* __restrict__ Vzo, const T* __restrict__ p,
                               const T* __restrict__ vx, const T* __restrict__ vy,
                               const T* __restrict__ vz, StepGeom g, ParamBox pb, T cv0, T cp0,
                               const T* __restrict__ g_row, T* gr_sum, T* gr_comp, T* gk_sum,
                               T* gk_comp) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz, n = sx * nx;
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int k = static_cast<int>(c % nz);
        const int j = static_cast<int>((c / nz) % ny);
        const int i = static_cast<int>(c / sx);
        // pb at a cell: decay product times (P + injected probe gradient)
        auto pbar = [&](int64_t cc, int ii, int jj, int kk) -> T {
            T v = P[cc];
            for (int q = 0; q < g.n_probe; ++q)
                if (g.probe_cell[q] == cc) v += g_row[q];
            const T lam = (static_cast<T>(g.dec[0][ii]) * static_cast<T>(g.dec[1][jj])) *
                          static_cast<T>(g.dec[2][kk]);
            return lam * v;
        };
        auto qv = [&](int64_t cc, int ii, int jj, int kk) -> T {
            if (!(g.flags[cc] & FDTD_FLAG_AIR)) return T(0);
            return cp0 * kappa_at<T>(pb, ii, jj, kk) * pbar(cc, ii, jj, kk);
        };
        const uint32_t f = g.flags[c];
        const T pbc = pbar(c, i, j, k);
        const T qc = (f & FDTD_FLAG_AIR) ? cp0 * kappa_at<T>(pb, i, j, k) * pbc : T(0);
        const T rc = r_at<T>(pb, i, j, k);
        const bool grad_here = in_box(pb, i, j, k);
        T acc_p = (f & FDTD_FLAG_AIR) ? pbc : T(0);
        T gr = T(0);
        // forward v' at the six faces of c (for the kappa gradient)
        T vxp = vx[c], vyp = vy[c], vzp = vz[c], vxmv = T(0), vymv = T(0), vzmv = T(0);
        const T pc = p[c];

        // ---- own faces (c = lo) ----
        {   // x
            const T d = static_cast<T>(g.dec[0][i]);
            if (i + 1 < nx) {
                const T rh = r_at<T>(pb, i + 1, j, k);
                const T cvf = cv0 / ((rc + rh) * T(0.5));
                const bool open = f & FDTD_FLAG_OPEN_X;
                const T w = open ? d * Vx[c] + qc - qv(c + sx, i + 1, j, k) : T(0);
                Vxo[c] = w;
                acc_p -= cvf * w;
                const T dp = p[c + sx] - pc;
                vxp = open ? vx[c] + cvf * dp : T(0);
                gr += w * dp * (-cvf / (rc + rh));
            } else {
                Vxo[c] = d * Vx[c] + qc;
            }
        }
        {   // y
            const T d = static_cast<T>(g.dec[1][j]);
            if (j + 1 < ny) {
                const T rh = r_at<T>(pb, i, j + 1, k);
                const T cvf = cv0 / ((rc + rh) * T(0.5));
                const bool open = f & FDTD_FLAG_OPEN_Y;
                const T w = open ? d * Vy[c] + qc - qv(c + sy, i, j + 1, k) : T(0);
                Vyo[c] = w;
                acc_p -= cvf * w;
                const T dp = p[c + sy] - pc;
                vyp = open ? vy[c] + cvf * dp : T(0);
                gr += w * dp * (-cvf / (rc + rh));
            } else {
                Vyo[c] = d * Vy[c] + qc;
            }
        }
        {   // z
            const T d = static_cast<T>(g.dec[2][k]);
            if (k + 1 < nz) {
                const T rh = r_at<T>(pb, i, j, k + 1);
                const T cvf = cv0 / ((rc + rh) * T(0.5));
                const bool open = f & FDTD_FLAG_OPEN_Z;
                const T w = open ? d * Vz[c] + qc - qv(c + 1, i, j, k + 1) : T(0);
                Vzo[c] = w;
                acc_p -= cvf * w;
                const T dp = p[c + 1] - pc;
                vzp = open ? vz[c] + cvf * dp : T(0);
                gr += w * dp * (-cvf / (rc + rh));
            } else {
                Vzo[c] = d * Vz[c] + qc;
            }
        }
        // ---- faces below (c = hi) ----
        if (i > 0) {
            const int64_t cl = c - sx;
            const T rl = r_at<T>(pb, i - 1, j, k);
            const T cvf = cv0 / ((rl + rc) * T(0.5));
            const bool open = g.flags[cl] & FDTD_FLAG_OPEN_X;
            const T w = open ? static_cast<T>(g.dec[0][i - 1]) * Vx[cl] + qv(cl, i - 1, j, k) - qc : T(0);
            acc_p += cvf * w;
            const T dp = pc - p[cl];
            vxmv = open ? vx[cl] + cvf * dp : T(0);
            gr += w * dp * (-cvf / (rl + rc));
        }
        if (j > 0) {
            const int64_t cl = c - sy;
            const T rl = r_at<T>(pb, i, j - 1, k);
            const T cvf = cv0 / ((rl + rc) * T(0.5));
            const bool open = g.flags[cl] & FDTD_FLAG_OPEN_Y;
            const T w = open ? static_cast<T>(g.dec[1][j - 1]) * Vy[cl] + qv(cl, i, j - 1, k) - qc : T(0);
            acc_p += cvf * w;
            const T dp = pc - p[cl];
            vymv = open ? vy[cl] + cvf * dp : T(0);
            gr += w * dp * (-cvf / (rl + rc));
        }
        if (k > 0) {
            const int64_t cl = c - 1;
            const T rl = r_at<T>(pb, i, j, k - 1);
            const T cvf = cv0 / ((rl + rc) * T(0.5));
            const bool open = g.flags[cl] & FDTD_FLAG_OPEN_Z;
            const T w = open ? static_cast<T>(g.dec[2][k - 1]) * Vz[cl] + qv(cl, i, j, k - 1) - qc : T(0);
            acc_p += cvf * w;
            const T dp = pc - p[cl];
            vzmv = open ? vz[cl] + cvf * dp : T(0);
            gr += w * dp * (-cvf / (rl + rc));
        }
        Po[c] = acc_p;
        if (grad_here && gr_sum) {
            const int64_t b = box_lin(pb, i, j, k);
            kahan_add(gr_sum, gr_comp, b, gr);
            if (f & FDTD_FLAG_AIR) {
                const T div = ((vxp - vxmv) + (vyp - vymv)) + (vzp - vzmv);
                kahan_add(gk_sum, gk_comp, b, pbc * cp0 * div);
            }
        }
    }
}

static inline int grid_cells(int64_t n) {
    int64_t b = (n + 255) / 256;
    return static_cast<int>(std::min<int64_t>(std::max<int64_t>(b, 1), 1 << 16));
}

template <typename T>
hipError_t launch_param_step(const T* p, const T* vx, const T* vy, const T* vz, T* po, T* vxo, T* vyo,
                             T* vzo, const StepGeom& g, const ParamBox& pb, T cv0, T cp0, double w,
                             T* trace_row, hipStream_t st) {
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(k_param_step<T>, dim3(grid_cells(n)), dim3(256), 0, st, p, vx, vy, vz, po, vxo,
                       vyo, vzo, g, pb, cv0, cp0, w, trace_row);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_adjoint_step(const T* P, const T* Vx, const T* Vy, const T* Vz, T* Po, T* Vxo, T* Vyo,
                               T* Vzo, const T* p, const T* vx, const T* vy, const T* vz,
                               const StepGeom& g, const ParamBox& pb, T cv0, T cp0, const T* g_row,
                               T* gr_sum, T* gr_comp, T* gk_sum, T* gk_comp, hipStream_t st) {
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(k_adjoint_step<T>, dim3(grid_cells(n)), dim3(256), 0, st, P, Vx, Vy, Vz, Po, Vxo,
                       Vyo, Vzo, p, vx, vy, vz, g, pb, cv0, cp0, g_row, gr_sum, gr_comp, gk_sum, gk_comp);
    return hipGetLastError();
}

}  // namespace fdtd_amd

using namespace fdtd_amd;

namespace {
StepGeom make_geom(int nx, int ny, int nz, const uint8_t* flags, const float* dx, const float* dy,
                   const float* dz, int64_t src_cell, int n_probe, const int64_t* probe_cell) {
    StepGeom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.flags = flags;
    g.dec[0] = dx; g.dec[1] = dy; g.dec[2] = dz;
    g.src_cell = src_cell;
    g.n_probe = n_probe;
    g.probe_cell = probe_cell;
    return g;
}
ParamBox make_box(const int32_t* lo, const int32_t* hi, const float* r, const float* kappa) {
    ParamBox b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = lo ? lo[a] : 0;
        b.hi[a] = hi ? hi[a] : 0;
    }
    b.r = r;
    b.kappa = kappa;
    return b;
}
}  // namespace

extern "C" {

int fdtd_param_step(int dtype, const void* p, const void* vx, const void* vy, const void* vz, void* po,
                    void* vxo, void* vyo, void* vzo, int nx, int ny, int nz, const uint8_t* flags,
                    const float* dx, const float* dy, const float* dz, const int32_t* box_lo,
                    const int32_t* box_hi, const float* r, const float* kappa, double cv0, double cp0,
                    int64_t src_cell, double w, int n_probe, const int64_t* probe_cell, void* trace_row,
                    void* stream) {
    if (nx < 2 || ny < 2 || nz < 2 || !flags || !dx || !dy || !dz) return -2;
    const StepGeom g = make_geom(nx, ny, nz, flags, dx, dy, dz, src_cell, n_probe, probe_cell);
    const ParamBox b = make_box(box_lo, box_hi, r, kappa);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (dtype == 8)
        e = launch_param_step<double>((const double*)p, (const double*)vx, (const double*)vy,
                                      (const double*)vz, (double*)po, (double*)vxo, (double*)vyo,
                                      (double*)vzo, g, b, cv0, cp0, w, (double*)trace_row, st);
    else
        e = launch_param_step<float>((const float*)p, (const float*)vx, (const float*)vy, (const float*)vz,
                                     (float*)po, (float*)vxo, (float*)vyo, (float*)vzo, g, b,
                                     static_cast<float>(cv0), static_cast<float>(cp0), w,
                                     (float*)trace_row, st);
    return e == hipSuccess ? 0 : -5;
}

int fdtd_adjoint_step(int dtype, const void* P, const void* Vx, const void* Vy, const void* Vz, void* Po,
                      void* Vxo, void* Vyo, void* Vzo, const void* p, const void* vx, const void* vy,
                      const void* vz, int nx, int ny, int nz, const uint8_t* flags, const float* dx,
                      const float* dy, const float* dz, const int32_t* box_lo, const int32_t* box_hi,
                      const float* r, const float* kappa, double cv0, double cp0, int n_probe,
                      const int64_t* probe_cell, const void* g_row, void* gr_sum, void* gr_comp,
                      void* gk_sum, void* gk_comp, void* stream) {
    if (nx < 2 || ny < 2 || nz < 2 || !flags || !dx || !dy || !dz) return -2;
    const StepGeom g = make_geom(nx, ny, nz, flags, dx, dy, dz, -1, n_probe, probe_cell);
    const ParamBox b = make_box(box_lo, box_hi, r, kappa);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (dtype == 8)
        e = launch_adjoint_step<double>(
            (const double*)P, (const double*)Vx, (const double*)Vy, (const double*)Vz, (double*)Po,
            (double*)Vxo, (double*)Vyo, (double*)Vzo, (const double*)p, (const double*)vx,
            (const double*)vy, (const double*)vz, g, b, cv0, cp0, (const double*)g_row, (double*)gr_sum,
            (double*)gr_comp, (double*)gk_sum, (double*)gk_comp, st);
    else
        e = launch_adjoint_step<float>(
            (const float*)P, (const float*)Vx, (const float*)Vy, (const float*)Vz, (float*)Po, (float*)Vxo,
            (float*)Vyo, (float*)Vzo, (const float*)p, (const float*)vx, (const float*)vy, (const float*)vz,
            g, b, static_cast<float>(cv0), static_cast<float>(cp0), (const float*)g_row, (float*)gr_sum,
            (float*)gr_comp, (float*)gk_sum, (float*)gk_comp, st);
    return e == hipSuccess ? 0 : -5;
}

// n_steps forward steps in one call. states: host array of 4 * n_states
// device pointers {p, vx, vy, vz}; pingpong != 0: step m reads states[m & 1]
// and writes states[(m + 1) & 1] (n_states = 2); otherwise step m reads
// states[m] and writes states[m + 1] (n_states = n_steps + 1). wave: host
// float64 waveform indexed by the absolute step; trace: device
// [steps][n_probe] (row step0 + m written by step m).
int fdtd_param_run(int dtype, void* const* states, int n_states, int64_t step0, int n_steps, int pingpong,
                   int nx, int ny, int nz, const uint8_t* flags, const float* dx, const float* dy,
                   const float* dz, const int32_t* box_lo, const int32_t* box_hi, const float* r,
                   const float* kappa, double cv0, double cp0, int64_t src_cell, const double* wave,
                   int n_probe, const int64_t* probe_cell, void* trace, void* stream) {
    if (!states || n_steps < 0 || !wave || (n_probe > 0 && (!trace || !probe_cell))) return -2;
    if (pingpong ? n_states != 2 : n_states != n_steps + 1) return -2;
    const size_t es = dtype == 8 ? 8 : 4;
    for (int m = 0; m < n_steps; ++m) {
        void* const* a = states + 4 * (pingpong ? (m & 1) : m);
        void* const* b = states + 4 * (pingpong ? ((m + 1) & 1) : m + 1);
        void* row = trace ? static_cast<char*>(trace) + (step0 + m) * n_probe * es : nullptr;
        if (int rc = fdtd_param_step(dtype, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], nx, ny, nz, flags,
                                     dx, dy, dz, box_lo, box_hi, r, kappa, cv0, cp0, src_cell,
                                     wave[step0 + m], n_probe, probe_cell, row, stream))
            return rc;
    }
    return 0;
}

// Adjoint steps n = n_hi - 1 down to n_lo. lam: host array of 8 device
// pointers, {P, Vx, Vy, Vz} of buffer 0 then buffer 1; buffer 0 holds
// lambda_{n_hi} on entry, and the result lambda_{n_lo} ends in buffer
// (n_hi - n_lo) & 1. states: host array of 4 * (n_hi - n_lo) pointers, the
// forward state of step n_lo + m at entry m. g: device [steps][n_probe]
// probe-trace gradients; acc: 4 device box arrays (Kahan sums).
int fdtd_adjoint_run(int dtype, void* const* lam, void* const* states, int64_t n_lo, int64_t n_hi, int nx,
                     int ny, int nz, const uint8_t* flags, const float* dx, const float* dy, const float* dz,
                     const int32_t* box_lo, const int32_t* box_hi, const float* r, const float* kappa,
                     double cv0, double cp0, int n_probe, const int64_t* probe_cell, const void* g,
                     void* gr_sum, void* gr_comp, void* gk_sum, void* gk_comp, void* stream) {
    if (!lam || !states || n_hi < n_lo || (n_probe > 0 && (!g || !probe_cell))) return -2;
    const size_t es = dtype == 8 ? 8 : 4;
    int cur = 0;
    for (int64_t n = n_hi - 1; n >= n_lo; --n) {
        void* const* a = lam + 4 * cur;
        void* const* b = lam + 4 * (cur ^ 1);
        void* const* st = states + 4 * (n - n_lo);
        const void* row = g ? static_cast<const char*>(g) + n * n_probe * es : nullptr;
        if (int rc = fdtd_adjoint_step(dtype, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], st[0], st[1],
                                       st[2], st[3], nx, ny, nz, flags, dx, dy, dz, box_lo, box_hi, r, kappa,
                                       cv0, cp0, n_probe, probe_cell, row, gr_sum, gr_comp, gk_sum, gk_comp,
                                       stream))
            return rc;
        cur ^= 1;
    }
    return 0;
}

}  // extern "C"
