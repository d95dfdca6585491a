// Fused FDTD time step for gfx950: velocity update, rigid-wall mask, pressure
// update, PML damping, source injection and probe gather in ONE sweep over HBM,
// plus the small kernels that add frequency-dependent (ADE) materials.
//
// Reference semantics (per cell, in this order; core/solver.py:2003-2193 with
// the native arithmetic of _kernels/src/fdtd_step.cpp, pml.cpp and ade.cpp):
//   v'   = v + cv*(p[c+e] - p[c])            faces below the last plane only
//   v'  += corr*(J[c+e] - J[c])               faces inside one pole material
//   v'   = open(face) ? v' : 0                rigid: either cell solid -> 0
//   J_m  = a*J_m + b*div_ade(v')              modulus poles (native divergence)
//   p'   = air ? p + cp*((dvx+dvy)+dvz) (+ corr*J_m ...) : 0
//   v    = v' * decay_axis ;  p = ((p' * dX) * dY) * dZ
//   p    = float(double(p) + w(t_s))          sources (air cells only)
//   trace[s][q] = p[probe q]
//
// Launch mapping (the 2.5D march): a wave owns a 64-wide k segment of R
// consecutive rows (j) and marches `planes` planes along i, the slowest axis.
// Per plane it reads p(i+1) once (p(i) is the previous plane's p(i+1), held in
// registers), v/flags once, and writes p, v once: 33 B per cell-update plus
// small halos (j-1/j+R rows, k0-1/k0+64 columns, one plane at the strip start).
// k neighbours inside the segment move across lanes with shuffles. No LDS, no
// barriers: waves are independent, so occupancy alone hides HBM latency.
// Ping-pong buffers (read set s&1, write the other) make the single sweep
// race-free.
//
// ADE materials: a face is corrected only when both of its cells hold the same
// pole material, and only material cells get a modulus correction, so only
// material cells' outputs differ from the plain update. A step with poles is
// therefore: the density-pole update over the material box (k_ade_density),
// the plain fused sweep over the whole grid, then k_ade_fixup, which recomputes
// the complete ADE update of every material cell from the (untouched) input
// buffers and overwrites its outputs.

#include <algorithm>

#include "fdtd_common.hpp"

namespace fdtd_amd {

__device__ __forceinline__ float shfl_down1(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, kWave); }

// Source q's add at cell (i, j, k) of step s: numpy adds the float64 sample
// to the float32 element in float64 and stores float32.
__device__ __forceinline__ float add_source(const FusedArgs& a, const GMEM double* __restrict__ WAVE, int q, int i,
                                            int j, int k, int64_t s, float ps) {
    bool on;
    if (a.src_kind[q] == 0) {
        on = a.src_i[q] == i && a.src_j[q] == j && a.src_k[q] == k;
    } else {
        const int ax = a.src_i[q], id = a.src_j[q];
        on = ax == 0 ? id == i : (ax == 1 ? id == j : id == k);
    }
    if (on) ps = static_cast<float>(static_cast<double>(ps) + WAVE[q * a.wave_len + s]);
    return ps;
}

// Probe q records the cell's pressure if it sits at (i, j, k).
__device__ __forceinline__ void record_probe(const FusedArgs& a, const GMEM int32_t* __restrict__ PIJK,
                                             GMEM float* __restrict__ TRACE, int q, int i, int j, int k, int64_t s,
                                             float ps) {
    if (PIJK[3 * q] == i && PIJK[3 * q + 1] == j && PIJK[3 * q + 2] == k) {
        TRACE[s * a.n_probe + q] = ps;
        if (!isfinite(ps)) atomicOr(a.health, 1u);
    }
}

// One wave's work: the strip (segment seg, j-strip sj, i-chunk si) of the
// solve described by `a` (kernel argument, or a row of a device batch table).
//
// W > 1 (stacked): the W waves of a workgroup own W consecutive j-strips of the
// same segment and i-chunk and march in step; per plane each wave posts its
// bottom/top rows (p, and vy/flags of the top row) to LDS, so the j halos
// between them come from LDS and only the workgroup's outer rows are read
// twice from HBM. `lds` holds 2 (plane parity) x 4 x W x 64 floats; `wi` is
// the wave's index in the stack. W == 1: every wave reads its own halos.
//
// NTS: store the step outputs with non-temporal hints. They are not read again
// within the step, and on a grid much larger than the caches this keeps L2 for
// the halo rows and columns the neighbouring strips re-read (about 5% faster at
// 512^3; about 3% on a batch of solves that together far exceed the caches).
// A working set that fits the Infinity Cache (kCacheResidentCells) is written
// with ordinary stores instead, so the next step finds its inputs on chip.
//
// IX: cell offset type, uint32_t for grids below 2^30 cells (scalar base +
// 32-bit offset addressing), int64_t beyond. LP: the solo sweeps' probe and
// source lookups (see below); the batch kernel, at its register cap, keeps the
// plain loops.
template <int R, bool NU, int W, bool NTS, typename IX, bool LP>
__device__ __forceinline__ void fused_strip(const FusedArgs& a, int64_t s, int seg, int sj, int si, int lane,
                                            float* lds, int wi) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave;
    const int k = k0 + lane;
    const bool kv = k < nz;
    const int j0 = sj * R;
    const int i0 = si * a.planes;
    const int i1 = min(i0 + a.planes, nx);
    const IX sx = static_cast<IX>(ny) * static_cast<IX>(nz);

    // s: index of the step this launch takes
    const int par = static_cast<int>(s & 1);
    // Every buffer is device (global) memory. Saying so lets the compiler use
    // global loads with a scalar base even when the arguments come from a
    // batch table in memory (otherwise it must assume flat addresses).
    const GMEM float* __restrict__ P = gmem(par ? a.f[1][0] : a.f[0][0]);
    const GMEM float* __restrict__ VX = gmem(par ? a.f[1][1] : a.f[0][1]);
    const GMEM float* __restrict__ VY = gmem(par ? a.f[1][2] : a.f[0][2]);
    const GMEM float* __restrict__ VZ = gmem(par ? a.f[1][3] : a.f[0][3]);
    GMEM float* __restrict__ Po = gmem(par ? a.f[0][0] : a.f[1][0]);
    GMEM float* __restrict__ VXo = gmem(par ? a.f[0][1] : a.f[1][1]);
    GMEM float* __restrict__ VYo = gmem(par ? a.f[0][2] : a.f[1][2]);
    GMEM float* __restrict__ VZo = gmem(par ? a.f[0][3] : a.f[1][3]);
    const GMEM uint8_t* __restrict__ FL = gmem(a.flags);
    const GMEM float* __restrict__ dX = gmem(a.dec[0]);
    const GMEM float* __restrict__ dY = gmem(a.dec[1]);
    const GMEM float* __restrict__ dZ = gmem(a.dec[2]);
    const GMEM float* __restrict__ ifc0 = gmem(a.ifc[0]);
    const GMEM float* __restrict__ ifc1 = gmem(a.ifc[1]);
    const GMEM float* __restrict__ ifc2 = gmem(a.ifc[2]);
    const GMEM float* __restrict__ icl0 = gmem(a.icl[0]);
    const GMEM float* __restrict__ icl1 = gmem(a.icl[1]);
    const GMEM float* __restrict__ icl2 = gmem(a.icl[2]);
    const GMEM double* __restrict__ WAVE = gmem(a.wave);
    const GMEM int32_t* __restrict__ PIJK = gmem(a.probe_ijk);
    GMEM float* __restrict__ TRACE = gmem(a.trace);

    // Which sources / probes can fall inside this strip (scalar work, once).
    uint32_t src_mask = 0;
    for (int q = 0; q < a.n_src; ++q) {
        bool hit;
        if (a.src_kind[q] == 0) {
            hit = a.src_i[q] >= i0 && a.src_i[q] < i1 && a.src_j[q] >= j0 && a.src_j[q] < j0 + R &&
                  a.src_k[q] >= k0 && a.src_k[q] < k0 + kWave;
        } else {
            const int ax = a.src_i[q], id = a.src_j[q];
            hit = ax == 0 ? (id >= i0 && id < i1)
                          : (ax == 1 ? (id >= j0 && id < j0 + R) : (id >= k0 && id < k0 + kWave));
        }
        if (hit) src_mask |= 1u << q;
    }
    // Probes in the strip. LP: lane q tests probe q (at most 64), so the
    // coordinates arrive in one round of vector loads rather than a chain of
    // scalar loads every wave waits on. The batch kernel, at its register
    // cap, keeps the scalar loop (its solves have few probes).
    uint64_t probe_mask = 0;
    if constexpr (LP) {
        bool probe_hit = false;
        if (lane < a.n_sweep_probe) {
            const int pi = PIJK[3 * lane], pj = PIJK[3 * lane + 1], pk = PIJK[3 * lane + 2];
            probe_hit = pi >= i0 && pi < i1 && pj >= j0 && pj < j0 + R && pk >= k0 && pk < k0 + kWave;
        }
        probe_mask = __builtin_amdgcn_ballot_w64(probe_hit);
    } else {
        for (int q = 0; q < a.n_sweep_probe; ++q) {
            const int pi = PIJK[3 * q], pj = PIJK[3 * q + 1], pk = PIJK[3 * q + 2];
            if (pi >= i0 && pi < i1 && pj >= j0 && pj < j0 + R && pk >= k0 && pk < k0 + kWave)
                probe_mask |= 1ull << q;
        }
    }
    const bool sources_ok = s < a.wave_len;
    const bool trace_ok = s < a.trace_len;
    if ((src_mask && !sources_ok) || (probe_mask && !trace_ok)) {
        if (lane == 0) atomicOr(a.health, 2u);
    }

    const float cv = a.cv, cp = a.cp;
    const float ifz_k = (NU && kv && k < nz - 1) ? ifc2[k] : 1.0f;
    const float icz_k = (NU && kv) ? icl2[k] : 1.0f;
    const float cvz = NU ? cv * ifz_k : cv;
    const float dz_k = kv ? dZ[k] : 1.0f;

    float pc[R], pn[R], vxm[R];
    bool rv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rv[r] = (j0 + r) < ny;
    // the rows' j tables, read once per strip rather than once per plane
    float dyr[R], cvyr[R], icyr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int jr = j0 + r;
        dyr[r] = rv[r] ? dY[jr] : 1.0f;
        cvyr[r] = (NU && jr + 1 < ny) ? cv * ifc1[jr] : cv;
        icyr[r] = (NU && rv[r]) ? icl1[jr] : 1.0f;
    }

    // ---- prologue: p(i0) and the x-face below the strip, vx'(i0-1) -------------
#pragma unroll
    for (int r = 0; r < R; ++r) {
        pc[r] = 0.f;
        vxm[r] = 0.f;
        if (rv[r] && kv) pc[r] = ldx(P, lidx<IX>(i0, j0 + r, k, ny, nz));
    }
    if (i0 > 0) {
        const int im = i0 - 1;
        const float cvx = NU ? cv * ifc0[im] : cv;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!(rv[r] && kv)) continue;
            const IX c = lidx<IX>(im, j0 + r, k, ny, nz);
            const float pm = ldx(P, c), vo = ldx(VX, c);
            const uint32_t f = ldx(FL, c);
            const float v = vo + cvx * (pc[r] - pm);
            vxm[r] = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
        }
    }

    // ---- march along i ---------------------------------------------------------
    for (int i = i0; i < i1; ++i) {
        const bool has_next = (i + 1) < nx;
        const float dx_i = dX[i];
        const float cvx = NU ? cv * (has_next ? ifc0[i] : 1.f) : cv;
        const float icx_i = NU ? icl0[i] : 1.f;

        float vxo[R], vyo[R], vzo[R];
        uint32_t fl[R];
        float hl_p[R], hl_vz[R], hr_p[R];
        uint32_t hl_f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pn[r] = 0.f;
            vxo[r] = vyo[r] = vzo[r] = 0.f;
            fl[r] = 0;
            hl_p[r] = hl_vz[r] = hr_p[r] = 0.f;
            hl_f[r] = 0;
            if (!rv[r]) continue;
            const IX c = lidx<IX>(i, j0 + r, k, ny, nz);
            if (kv) {
                if (has_next) pn[r] = ldx(P, c + sx);
                vxo[r] = ldx(VX, c);
                vyo[r] = ldx(VY, c);
                vzo[r] = ldx(VZ, c);
                fl[r] = ldx(FL, c);
            }
            if (lane == 0 && k0 > 0) {
                hl_p[r] = ldx(P, c - 1);
                hl_vz[r] = ldx(VZ, c - 1);
                hl_f[r] = ldx(FL, c - 1);
            }
            if (lane == kWave - 1 && k + 1 < nz) hr_p[r] = ldx(P, c + 1);
        }
        // j halos: p(j0-1), vy(j0-1), flags(j0-1) and p(j0+R)
        float p_jm = 0.f, vy_jm = 0.f, p_jp = 0.f;
        uint32_t f_jm = 0;
        if constexpr (W > 1) {
            float* sb = lds + (i & 1) * (4 * W * kWave);
            sb[(0 * W + wi) * kWave + lane] = pc[0];
            sb[(1 * W + wi) * kWave + lane] = pc[R - 1];
            sb[(2 * W + wi) * kWave + lane] = vyo[R - 1];
            sb[(3 * W + wi) * kWave + lane] = __uint_as_float(fl[R - 1]);
            __syncthreads();
            if (wi > 0) {
                p_jm = sb[(1 * W + wi - 1) * kWave + lane];
                vy_jm = sb[(2 * W + wi - 1) * kWave + lane];
                f_jm = __float_as_uint(sb[(3 * W + wi - 1) * kWave + lane]);
            }
            if (wi < W - 1) p_jp = sb[(0 * W + wi + 1) * kWave + lane];
        }
        if ((W == 1 || wi == 0) && kv && j0 > 0) {
            const IX c = lidx<IX>(i, j0 - 1, k, ny, nz);
            p_jm = ldx(P, c);
            vy_jm = ldx(VY, c);
            f_jm = ldx(FL, c);
        }
        if ((W == 1 || wi == W - 1) && kv && j0 + R < ny) p_jp = ldx(P, lidx<IX>(i, j0 + R, k, ny, nz));

        // vy' on the face (j0-1 | j0)
        float vyprev = 0.f;
        if (j0 > 0) {
            const float cvy = NU ? cv * ifc1[j0 - 1] : cv;
            const float v = vy_jm + cvy * (pc[0] - p_jm);
            vyprev = (f_jm & FDTD_FLAG_OPEN_Y) ? v : 0.f;
        }

#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rv[r]) continue;  // wave-uniform
            const int jr = j0 + r;
            const IX c = lidx<IX>(i, jr, k, ny, nz);
            const float pr = pc[r];
            const float p_up = (r + 1 < R) ? pc[r + 1] : p_jp;
            const uint32_t f = fl[r];

            // x face (i | i+1)
            float vxn = vxo[r];
            if (has_next) {
                const float v = vxo[r] + cvx * (pn[r] - pr);
                vxn = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
            // y face (j | j+1)
            float vyn = vyo[r];
            if (jr + 1 < ny) {
                const float v = vyo[r] + cvyr[r] * (p_up - pr);
                vyn = (f & FDTD_FLAG_OPEN_Y) ? v : 0.f;
            }
            // z face (k | k+1): p(k+1) from the next lane, or the right halo
            float pk1 = shfl_down1(pr);
            if (lane == kWave - 1) pk1 = hr_p[r];
            float vzn = vzo[r];
            if (k + 1 < nz) {
                const float v = vzo[r] + cvz * (pk1 - pr);
                vzn = (f & FDTD_FLAG_OPEN_Z) ? v : 0.f;
            }
            // z face (k-1 | k): previous lane, or the left halo for lane 0
            float vzm = shfl_up1(vzn);
            if (lane == 0) {
                vzm = 0.f;
                if (k0 > 0) {
                    const float cz = NU ? cv * ifc2[k0 - 1] : cv;
                    const float v = hl_vz[r] + cz * (pr - hl_p[r]);
                    vzm = (hl_f[r] & FDTD_FLAG_OPEN_Z) ? v : 0.f;
                }
            }

            // pressure
            const float dvx = vxn - vxm[r];
            const float dvy = vyn - vyprev;
            const float dvz = vzn - vzm;
            float pnew;
            if (NU) {
                const float ddx = icx_i * dvx, ddy = icyr[r] * dvy, ddz = icz_k * dvz;
                pnew = pr + cp * ((ddx + ddy) + ddz);
            } else {
                pnew = pr + cp * ((dvx + dvy) + dvz);
            }
            if (!(f & FDTD_FLAG_AIR)) pnew = 0.f;

            // PML (interior decay is exactly 1.0f, so the multiply is exact there)
            const float dy_j = dyr[r];
            float ps = ((pnew * dx_i) * dy_j) * dz_k;
            if (src_mask && sources_ok && (f & FDTD_FLAG_AIR)) {
                // in list order: the set bits (LP), or every source with a test
                if constexpr (LP) {
                    for (uint32_t sm = src_mask; sm; sm &= sm - 1)
                        ps = add_source(a, WAVE, __builtin_ctz(sm), i, jr, k, s, ps);
                } else {
                    for (int q = 0; q < a.n_src; ++q)
                        if ((src_mask >> q) & 1u) ps = add_source(a, WAVE, q, i, jr, k, s, ps);
                }
            }
            if (kv) {
                if constexpr (NTS) {
                    stx_nt(Po, c, ps);
                    stx_nt(VXo, c, vxn * dx_i);
                    stx_nt(VYo, c, vyn * dy_j);
                    stx_nt(VZo, c, vzn * dz_k);
                } else {
                    stx(Po, c, ps);
                    stx(VXo, c, vxn * dx_i);
                    stx(VYo, c, vyn * dy_j);
                    stx(VZo, c, vzn * dz_k);
                }
            }
            if (probe_mask && trace_ok) {
                if constexpr (LP) {
                    for (uint64_t pm = probe_mask; pm; pm &= pm - 1)
                        record_probe(a, PIJK, TRACE, __builtin_ctzll(pm), i, jr, k, s, ps);
                } else {
                    for (int q = 0; q < a.n_probe; ++q)
                        if ((probe_mask >> q) & 1ull) record_probe(a, PIJK, TRACE, q, i, jr, k, s, ps);
                }
            }
            vxm[r] = vxn;
            vyprev = vyn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) pc[r] = pn[r];
    }
}

__device__ __forceinline__ void ade_density_part(const FusedArgs& a, int64_t first, int64_t stride);

// The nonuniform 4-row sweep with 32-bit offsets is held to 72 VGPRs (7 waves
// per SIMD instead of 6, no spills): 1.1% faster on a 320x320x480 grid.
// Blocks past a.sweep_blocks run the step's ADE density update: it reads the
// input pressure and writes only J, which the sweep never touches, so it
// fills the sweep's tail instead of taking a launch of its own.
template <int R, bool NU, typename IX, bool NTS>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu((NU && R == 4 && sizeof(IX) == 4) ? 7 : 1, 8))) void k_fused_step(
    const FusedArgs a) {
    if (a.dens_blocks > 0 && blockIdx.x >= static_cast<unsigned>(a.sweep_blocks)) {
        ade_density_part(a, static_cast<int64_t>(blockIdx.x - a.sweep_blocks) * kBlock + threadIdx.x,
                         static_cast<int64_t>(a.dens_blocks) * kBlock);
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    // XCD-aware remap: hardware block b runs on XCD b % 8; give each XCD a
    // contiguous range of logical blocks so strips that share halo rows share
    // an L2. gridDim.x is a multiple of 8.
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    // wave index within the block, made provably wave-uniform (scalar registers)
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= a.nwaves) return;
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;
    const int seg = w % a.nseg;
    const int t = w / a.nseg;
    fused_strip<R, NU, 1, NTS, IX, true>(a, s, seg, t % a.nstrip_j, t / a.nstrip_j, lane, nullptr, 0);
}

// Stacked variant: workgroup = W waves over W consecutive j-strips of one
// segment and i-chunk (halo rows shared through LDS). Logical workgroup
// index -> (segment, j-stack, i-chunk), segment fastest; `a.nwaves` counts
// workgroups here. Every wave of a workgroup runs the same number of planes,
// so the per-plane barrier is reached uniformly. Grids below 2^30 cells.
template <int R, bool NU, int W>
__global__ __launch_bounds__(W * kWave) void k_fused_stack(const FusedArgs a) {
    __shared__ float lds[2 * 4 * W * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wi = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    if (lb >= a.nwaves) return;                                  // whole workgroup
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;                               // whole workgroup
    const int nstack_j = (a.nstrip_j + W - 1) / W;
    const int seg = lb % a.nseg;
    const int t = lb / a.nseg;
    fused_strip<R, NU, W, true, uint32_t, true>(a, s, seg, (t % nstack_j) * W + wi, t / nstack_j, lane, lds, wi);
}

// Batched step: one launch advances every solve in `table` by one step, step
// `koff` of a chunk whose first step is each solve's counter (the counters
// move once per chunk, by k_batch_advance). Waves are dealt out by solve
// through the prefix sums `offsets` (n + 1 entries); a solve whose step
// reached its limit is left untouched. Solves below 2^30 cells.
// The number of solves is read from device memory (n_ptr), so a captured
// graph stays valid while the host rewrites the table between chunks; the
// grid is sized for the table's wave capacity and surplus waves return.
// Register budget: at most 80 VGPRs (6 waves per SIMD) rather than the 83 the
// compiler picks unguided; measured about 8% faster on the 64-request sweep.
template <int R, bool NTS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_batch_step(
    const FusedArgs* __restrict__ table, const int* __restrict__ offsets, const int* __restrict__ n_ptr, int koff) {
    const int lane = threadIdx.x & (kWave - 1);
    // XCD-aware remap over the blocks the live table needs (the grid may be
    // larger): XCD x takes logical blocks [x * bpx, (x + 1) * bpx)
    const int n = *n_ptr;
    const int total = offsets[n];
    const int bpx = ((total + kBlockWaves - 1) / kBlockWaves + 7) / 8;
    const int xb = blockIdx.x >> 3;
    if (xb >= bpx) return;
    const int lb = (blockIdx.x & 7) * bpx + xb;
    const int gw = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (gw >= total) return;
    int lo = 0, hi = n;  // offsets[lo] <= gw < offsets[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (offsets[mid] <= gw) lo = mid; else hi = mid;
    }
    const FusedArgs& a = table[lo];
    const int w = gw - offsets[lo];
    const int64_t s = *a.step_ptr + koff;
    if (s >= a.step_limit) return;
    const int seg = w % a.nseg;
    const int t = w / a.nseg;
    fused_strip<R, false, 1, NTS, uint32_t, false>(a, s, seg, t % a.nstrip_j, t / a.nstrip_j, lane, nullptr, 0);
}

// End of a chunk of n batched steps: every solve's counter moves by n, up to
// its limit.
__global__ void k_batch_advance(const FusedArgs* __restrict__ table, const int* __restrict__ n_ptr, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *n_ptr) return;
    int64_t* step = const_cast<int64_t*>(table[e].step_ptr);
    const int64_t v = step[0], lim = table[e].step_limit;
    if (v < lim) step[0] = v + n < lim ? v + n : lim;
}

// ---- ADE --------------------------------------------------------------------
__device__ __forceinline__ int64_t box_index(const FusedArgs& a, int i, int j, int k) {
    const int by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
    return ((static_cast<int64_t>(i - a.box_lo[0]) * by) + (j - a.box_lo[1])) * bz + (k - a.box_lo[2]);
}

__device__ __forceinline__ int64_t box_cells(const FusedArgs& a) {
    return static_cast<int64_t>(a.box_hi[0] - a.box_lo[0]) * (a.box_hi[1] - a.box_lo[1]) *
           (a.box_hi[2] - a.box_lo[2]);
}

// Density-pole auxiliary update over the ADE box, run before the sweep of the
// same time step (J <- a J + b p, Lorentz with history), ade.cpp:25-166.
__device__ __forceinline__ void ade_density_part(const FusedArgs& a, int64_t first, int64_t stride) {
    const int by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
    const int64_t box_n = box_cells(a);
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;
    const float* __restrict__ P = (s & 1) ? a.f[1][0] : a.f[0][0];
    const int64_t n_visit = a.ade_cells ? a.n_ade_cells : box_n;
    for (int64_t e = first; e < n_visit; e += stride) {
        const int64_t b = a.ade_cells ? static_cast<int64_t>(a.ade_cells[e]) : e;
        const int kk = static_cast<int>(b % bz) + a.box_lo[2];
        const int jj = static_cast<int>((b / bz) % by) + a.box_lo[1];
        const int ii = static_cast<int>(b / (static_cast<int64_t>(by) * bz)) + a.box_lo[0];
        const int64_t c = lin(ii, jj, kk, a.ny, a.nz);
        const int m = a.material[c];
        if (!m) continue;
        const float pv = P[c];
        for (int q = 0; q < a.n_pole; ++q) {
            const AdePole pl = pole_at(a.poles, q);
            if (pl.target != 0 || pl.mat != m) continue;
            float* Jq = a.J + q * box_n;
            if (!pl.lorentz) {
                Jq[b] = pl.a * Jq[b] + pl.b * pv;
            } else {
                float* Jp = a.Jprev + q * box_n;
                const float j0v = Jq[b], j1v = Jp[b];
                const float jn = pl.a * j0v + pl.b * j1v + pl.d * pv;
                Jp[b] = j0v;
                Jq[b] = jn;
            }
        }
    }
}

__global__ void k_ade_density(const FusedArgs a) {
    ade_density_part(a, blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x,
                     static_cast<int64_t>(gridDim.x) * blockDim.x);
}

// v' of the face between cells lo (flat c) and hi (c + step along `axis`), with
// the density-pole correction (Debye poles first, then Lorentz, ade.cpp:240-401).
template <bool NU>
__device__ __forceinline__ float ade_face(const FusedArgs& a, const float* __restrict__ P,
                                          const float* __restrict__ V, int axis, int i, int j, int k,
                                          int64_t c, int64_t step) {
    const int n_ax = axis == 0 ? a.nx : (axis == 1 ? a.ny : a.nz);
    const int idx = axis == 0 ? i : (axis == 1 ? j : k);
    const float vo = V[c];
    if (idx + 1 >= n_ax) return vo;  // last face: not updated, not masked
    const float cf = NU ? a.cv * a.ifc[axis][idx] : a.cv;
    float v = vo + cf * (P[c + step] - P[c]);
    const int m_lo = a.material[c], m_hi = a.material[c + step];
    if (m_lo && m_lo == m_hi) {
        const int64_t box_n = box_cells(a);
        const int ih = i + (axis == 0), jh = j + (axis == 1), kh = k + (axis == 2);
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < a.n_pole; ++q) {
                const AdePole pl = pole_at(a.poles, q);
                if (pl.target != 0 || pl.lorentz != pass || pl.mat != m_lo) continue;
                const float* Jq = a.J + q * box_n;
                v += pl.corr * (Jq[box_index(a, ih, jh, kh)] - Jq[box_index(a, i, j, k)]);
            }
    }
    const uint32_t open = axis == 0 ? FDTD_FLAG_OPEN_X : (axis == 1 ? FDTD_FLAG_OPEN_Y : FDTD_FLAG_OPEN_Z);
    return (a.flags[c] & open) ? v : 0.f;
}

// Full ADE update of every pole-material cell in the box, from the input set,
// overwriting what the plain sweep wrote for those cells (see file comment).
template <bool NU>
__global__ void k_ade_fixup(const FusedArgs a) {
    const int ny = a.ny, nz = a.nz;
    const int by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
    const int64_t box_n = box_cells(a);
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;
    const int par = static_cast<int>(s & 1);
    const float* __restrict__ P = par ? a.f[1][0] : a.f[0][0];
    const float* __restrict__ VX = par ? a.f[1][1] : a.f[0][1];
    const float* __restrict__ VY = par ? a.f[1][2] : a.f[0][2];
    const float* __restrict__ VZ = par ? a.f[1][3] : a.f[0][3];
    float* __restrict__ Po = par ? a.f[0][0] : a.f[1][0];
    float* __restrict__ VXo = par ? a.f[0][1] : a.f[1][1];
    float* __restrict__ VYo = par ? a.f[0][2] : a.f[1][2];
    float* __restrict__ VZo = par ? a.f[0][3] : a.f[1][3];
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
    const int64_t n_visit = a.ade_cells ? a.n_ade_cells : box_n;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n_visit;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = a.ade_cells ? static_cast<int64_t>(a.ade_cells[e]) : e;
        const int k = static_cast<int>(b % bz) + a.box_lo[2];
        const int j = static_cast<int>((b / bz) % by) + a.box_lo[1];
        const int i = static_cast<int>(b / (static_cast<int64_t>(by) * bz)) + a.box_lo[0];
        const int64_t c = lin(i, j, k, ny, nz);
        const int m = a.material[c];
        bool has_pole = false;
        for (int q = 0; q < a.n_pole; ++q) has_pole |= pole_at(a.poles, q).mat == m;
        if (!m || !has_pole) continue;
        const uint32_t f = a.flags[c];
        const float vxn = ade_face<NU>(a, P, VX, 0, i, j, k, c, sx);
        const float vyn = ade_face<NU>(a, P, VY, 1, i, j, k, c, sy);
        const float vzn = ade_face<NU>(a, P, VZ, 2, i, j, k, c, 1);
        const float vxm = i > 0 ? ade_face<NU>(a, P, VX, 0, i - 1, j, k, c - sx, sx) : 0.f;
        const float vym = j > 0 ? ade_face<NU>(a, P, VY, 1, i, j - 1, k, c - sy, sy) : 0.f;
        const float vzm = k > 0 ? ade_face<NU>(a, P, VZ, 2, i, j, k - 1, c - 1, 1) : 0.f;
        const float dvx = vxn - vxm, dvy = vyn - vym, dvz = vzn - vzm;
        const float icx = NU ? a.icl[0][i] : 1.f, icy = NU ? a.icl[1][j] : 1.f, icz = NU ? a.icl[2][k] : 1.f;
        // modulus poles: source is the native divergence of v'
        const float ix = NU ? icx : a.ade_inv_dx, iy = NU ? icy : a.ade_inv_dx, iz = NU ? icz : a.ade_inv_dx;
        float dA = dvx * ix;
        dA += dvy * iy;
        dA += dvz * iz;
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < a.n_pole; ++q) {
                const AdePole pl = pole_at(a.poles, q);
                if (pl.target != 1 || pl.lorentz != pass || pl.mat != m) continue;
                float* Jq = a.J + q * box_n;
                if (!pass) {
                    Jq[b] = pl.a * Jq[b] + pl.b * dA;
                } else {
                    float* Jp = a.Jprev + q * box_n;
                    const float j0v = Jq[b], j1v = Jp[b];
                    const float jn = pl.a * j0v + pl.b * j1v + pl.d * dA;
                    Jp[b] = j0v;
                    Jq[b] = jn;
                }
            }
        float pnew;
        if (NU) pnew = P[c] + a.cp * (((icx * dvx) + (icy * dvy)) + (icz * dvz));
        else pnew = P[c] + a.cp * ((dvx + dvy) + dvz);
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < a.n_pole; ++q) {
                const AdePole pl = pole_at(a.poles, q);
                if (pl.target != 1 || pl.lorentz != pass || pl.mat != m) continue;
                pnew += pl.corr * a.J[q * box_n + b];
            }
        if (!(f & FDTD_FLAG_AIR)) pnew = 0.f;
        const float dx = a.dec[0][i], dy = a.dec[1][j], dz = a.dec[2][k];
        float ps = ((pnew * dx) * dy) * dz;
        if (s < a.wave_len && (f & FDTD_FLAG_AIR)) {
            for (int q = 0; q < a.n_src; ++q) {
                bool on;
                if (a.src_kind[q] == 0) {
                    on = a.src_i[q] == i && a.src_j[q] == j && a.src_k[q] == k;
                } else {
                    const int ax = a.src_i[q], id = a.src_j[q];
                    on = ax == 0 ? id == i : (ax == 1 ? id == j : id == k);
                }
                if (on) ps = static_cast<float>(static_cast<double>(ps) + a.wave[q * a.wave_len + s]);
            }
        }
        Po[c] = ps;
        VXo[c] = vxn * dx;
        VYo[c] = vyn * dy;
        VZo[c] = vzn * dz;
        if (s < a.trace_len)
            for (int q = 0; q < a.n_sweep_probe; ++q)
                if (a.probe_ijk[3 * q] == i && a.probe_ijk[3 * q + 1] == j && a.probe_ijk[3 * q + 2] == k) {
                    a.trace[s * a.n_probe + q] = ps;
                    if (!isfinite(ps)) atomicOr(a.health, 1u);
                }
    }
}

// Microphones: after step s the state is in set (s+1)&1; each thread gathers
// one microphone's p, vx, vy, vz from 8 corners each in the corner order of
// core/solver.py:962-1003 (acc = 0, then acc += w * f). The reference sums in
// float32 when its weights are Python floats (numpy rounds the weight to
// float32 and multiplies in float32) and in float64 when they are numpy
// float64 scalars; mic_f64 selects which, per gather. Results are stored as
// float64 (a float32 sum exactly).
__global__ void k_mic_gather(const FusedArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.n_mic) return;
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit || s >= a.mic_len) return;
    const int par = static_cast<int>((s + 1) & 1);
    for (int g = 0; g < 4; ++g) {
        const float* __restrict__ F = a.f[par][g];
        const int64_t row = static_cast<int64_t>(m) * 4 + g;
        const int32_t* idx = a.mic_idx + row * 8;
        const double* w = a.mic_w + row * 8;
        double out;
        if (a.mic_f64[row]) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 8; ++c) acc = acc + w[c] * static_cast<double>(F[idx[c]]);
            out = acc;
        } else {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) acc = acc + static_cast<float>(w[c]) * F[idx[c]];
            out = acc;
        }
        a.mic_out[(s * a.n_mic + m) * 4 + g] = out;
    }
}

// Probes n_sweep_probe.. n_probe-1 of step s (the sweep records the first
// n_sweep_probe): after every pass of the step has written the output set
// (s+1)&1, one thread per probe copies p at its cell into the trace, as
// _record_probes does after all sources (core/solver.py:2435-2439).
__global__ void k_probe_tail(const FusedArgs a) {
    const int q = a.n_sweep_probe + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= a.n_probe) return;
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;
    if (s >= a.trace_len) {
        atomicOr(a.health, 2u);
        return;
    }
    const int par = static_cast<int>((s + 1) & 1);
    const int64_t c = lin(a.probe_ijk[3 * q], a.probe_ijk[3 * q + 1], a.probe_ijk[3 * q + 2], a.ny, a.nz);
    const float ps = a.f[par][0][c];
    a.trace[s * a.n_probe + q] = ps;
    if (!isfinite(ps)) atomicOr(a.health, 1u);
}

// Boundary plane `bp` of step s on the output pressure (s+1)&1, after the
// sweep and the ADE fix-up, before the table sources and the probes: one
// thread per cell of the plane's cross-section, which owns that column's
// cells on the faces (and reads its inner planes), in the reference's
// statement order (boundaries/_boundaries.py:479-511 Mur, :707-776 radiation).
// Float32 differences, float64 products and sums with coeff / omr, refl in
// float32, one rounding per stored value: numpy's types for a float64
// coefficient scalar and a Python-float R.
__global__ void k_boundary_plane(const FusedArgs a, const fdtd_boundary_plane bp, float* __restrict__ state) {
    const int64_t s = *a.step_ptr + a.koff;
    if (s >= a.step_limit) return;
    const int ax = bp.axis;
    const int N = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
    const int nu = ax == 0 ? a.ny : a.nx, nw = ax == 2 ? a.ny : a.nz;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(nu) * nw) return;
    const int u = static_cast<int>(t / nw), w = static_cast<int>(t % nw);
    float* __restrict__ P = a.f[(s + 1) & 1][0];
    auto cell = [&](int c) -> int64_t {
        return ax == 0 ? lin(c, u, w, a.ny, a.nz) : (ax == 1 ? lin(u, c, w, a.ny, a.nz) : lin(u, w, c, a.ny, a.nz));
    };
    float* st = state + bp.state_off;
    if (bp.kind == 0) {
        float* prev0 = st;
        float* prev1 = st + static_cast<int64_t>(nu) * nw;
        const int64_t c0 = cell(0), c1 = cell(1), cl = cell(N - 1), cm = cell(N - 2);
        const float d0 = P[c1] - P[c0];
        P[c0] = static_cast<float>(static_cast<double>(prev0[t]) + bp.coeff * static_cast<double>(d0));
        prev0[t] = P[c1];
        const float d1 = P[cm] - P[cl];
        P[cl] = static_cast<float>(static_cast<double>(prev1[t]) + bp.coeff * static_cast<double>(d1));
        prev1[t] = P[cm];
    } else {
        const int64_t cb = cell(bp.side ? N - 1 : 0), ci = cell(bp.side ? N - 2 : 1);
        const float pin = P[ci];
        const float d = pin - P[cb];
        const double abc = static_cast<double>(st[t]) + bp.coeff * static_cast<double>(d);
        const float rig = static_cast<float>(bp.refl) * pin;
        P[cb] = static_cast<float>(static_cast<double>(rig) + bp.omr * abc);
        st[t] = pin;
    }
}

// Membrane sources of step s (core/solver.py:2389-2412): after the sweep and
// the ADE fix-up the step's output is in set (s+1)&1. One thread per listed
// cell applies the cell's entries in source order,
//   field[c] = float(double(field[c]) + wave[src][s] * weight)
// (numpy adds the float64 product to the float32 element in float64 and
// stores float32), then records the probes on the cell, which the reference
// samples after every source.
__global__ void k_membrane(const FusedArgs a) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n_mcell) return;
    // The loads that do not depend on one another are issued together: with
    // one entry per cell (m_single) entry e belongs to cell e, so the entry
    // needs no lookup through m_estart.
    const int64_t s = *a.step_ptr + a.koff;
    const int64_t c = a.mcell[e];
    const int t0 = a.m_single ? e : a.m_estart[e];
    const int t1 = a.m_single ? e + 1 : a.m_estart[e + 1];
    if (s >= a.step_limit) return;
    if (s >= a.m_wave_len) {
        atomicOr(a.health, 2u);
        return;
    }
    const int out = static_cast<int>((s + 1) & 1);
    for (int t = t0; t < t1; ++t) {
        float* F = a.f[out][a.m_efield[t]];
        const double add = a.m_wave[static_cast<int64_t>(a.m_esrc[t]) * a.m_wave_len + s] * a.m_eweight[t];
        F[c] = static_cast<float>(static_cast<double>(F[c]) + add);
    }
    if (s < a.trace_len) {
        const float pv = a.f[out][0][c];
        for (int t = a.m_pstart[e]; t < a.m_pstart[e + 1]; ++t) {
            a.trace[s * a.n_probe + a.m_pidx[t]] = pv;
            if (!isfinite(pv)) atomicOr(a.health, 1u);
        }
    }
}

__global__ void k_advance(int64_t* step, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] = step[0] + n;
}

// ---- host-side launch helpers ------------------------------------------------
template <int R, bool NU>
static hipError_t launch_r(const FusedArgs& a, int blocks, hipStream_t st) {
    const int64_t cells = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    if (cells <= kCacheResidentCells)
        hipLaunchKernelGGL((k_fused_step<R, NU, uint32_t, false>), dim3(blocks), dim3(kBlock), 0, st, a);
    else if (cells < kMaxCells32)
        hipLaunchKernelGGL((k_fused_step<R, NU, uint32_t, true>), dim3(blocks), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((k_fused_step<R, NU, int64_t, true>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int R, bool NU, int W>
static hipError_t launch_stack(const FusedArgs& a, int blocks, hipStream_t st) {
    if (static_cast<int64_t>(a.nx) * a.ny * a.nz >= kMaxCells32) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fused_stack<R, NU, W>), dim3(blocks), dim3(W * kWave), 0, st, a);
    return hipGetLastError();
}

template <int R, bool NU>
static hipError_t launch_rn(const FusedArgs& a, int stack, int blocks, hipStream_t st) {
    switch (stack) {
        case 1: return launch_r<R, NU>(a, blocks, st);
        case 2: return launch_stack<R, NU, 2>(a, blocks, st);
        case 4: return launch_stack<R, NU, 4>(a, blocks, st);
        case 8: return launch_stack<R, NU, 8>(a, blocks, st);
        default: return hipErrorInvalidValue;
    }
}

template <int R>
static hipError_t launch_rows(const FusedArgs& a, bool nu, int stack, int blocks, hipStream_t st) {
    return nu ? launch_rn<R, true>(a, stack, blocks, st) : launch_rn<R, false>(a, stack, blocks, st);
}

// Fill the tiling fields of `a` for rows R / planes / stack; returns the grid
// size (a multiple of 8, the XCD count). stack == 1: one independent wave per
// strip, 4 waves per workgroup, `nwaves` = strips. stack > 1: one workgroup of
// `stack` waves per j-stack, `nwaves` = workgroups.
int fused_grid(FusedArgs& a, int rows, int planes, int stack) {
    a.planes = planes;
    a.nseg = (a.nz + kWave - 1) / kWave;
    a.nstrip_j = (a.ny + rows - 1) / rows;
    a.nstrip_i = (a.nx + planes - 1) / planes;
    int blocks;
    if (stack <= 1) {
        a.nwaves = a.nseg * a.nstrip_j * a.nstrip_i;
        blocks = (a.nwaves + kBlockWaves - 1) / kBlockWaves;
    } else {
        a.nwaves = a.nseg * ((a.nstrip_j + stack - 1) / stack) * a.nstrip_i;
        blocks = a.nwaves;
    }
    blocks = (blocks + 7) & ~7;
    a.blocks_per_xcd = blocks / 8;
    return blocks;
}

// One full step: [ADE density] -> fused sweep -> [ADE fix-up] -> [boundary
// planes] -> [table sources] -> [probes past the sweep's] -> [microphones].
// ev_before / ev_after (optional) are recorded around the sweep alone, for
// timing it. bnd: the context's boundary planes, or null.
hipError_t launch_fused(const FusedArgs& a_in, int rows, int stack, bool nu, bool ade, int blocks,
                        hipStream_t st, hipEvent_t ev_before, hipEvent_t ev_after, const BoundarySet* bnd) {
    FusedArgs a = a_in;
    a.sweep_blocks = blocks;
    a.dens_blocks = 0;
    int box_blocks = 1;
    if (ade) {
        const int64_t bn = a.ade_cells ? a.n_ade_cells
                                       : static_cast<int64_t>(a.box_hi[0] - a.box_lo[0]) *
                                             (a.box_hi[1] - a.box_lo[1]) * (a.box_hi[2] - a.box_lo[2]);
        box_blocks = static_cast<int>(std::min<int64_t>(std::max<int64_t>((bn + 255) / 256, 1), 1 << 16));
        if (stack == 1) {
            a.dens_blocks = box_blocks;  // run inside the sweep launch (k_fused_step)
        } else {
            hipLaunchKernelGGL(k_ade_density, dim3(box_blocks), dim3(256), 0, st, a);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    hipError_t e;
    if (ev_before && (e = hipEventRecord(ev_before, st)) != hipSuccess) return e;
    const int grid = blocks + a.dens_blocks;
    switch (rows) {
        case 2: e = launch_rows<2>(a, nu, stack, grid, st); break;
        case 4: e = launch_rows<4>(a, nu, stack, grid, st); break;
        case 8: e = launch_rows<8>(a, nu, stack, grid, st); break;
        case 16: e = launch_rows<16>(a, nu, stack, grid, st); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (ev_after && (e = hipEventRecord(ev_after, st)) != hipSuccess) return e;
    if (ade) {
        if (nu) hipLaunchKernelGGL(k_ade_fixup<true>, dim3(box_blocks), dim3(256), 0, st, a);
        else hipLaunchKernelGGL(k_ade_fixup<false>, dim3(box_blocks), dim3(256), 0, st, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (bnd) {
        for (int b = 0; b < bnd->n; ++b) {
            const fdtd_boundary_plane& bp = bnd->planes[b];
            const int64_t cross = static_cast<int64_t>(a.nx) * a.ny * a.nz /
                                  (bp.axis == 0 ? a.nx : (bp.axis == 1 ? a.ny : a.nz));
            hipLaunchKernelGGL(k_boundary_plane, dim3(static_cast<unsigned>((cross + 255) / 256)), dim3(256), 0, st,
                               a, bp, bnd->state);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    if (a.n_mcell > 0) {
        hipLaunchKernelGGL(k_membrane, dim3((a.n_mcell + 255) / 256), dim3(256), 0, st, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.n_probe > a.n_sweep_probe) {
        hipLaunchKernelGGL(k_probe_tail, dim3((a.n_probe - a.n_sweep_probe + 255) / 256), dim3(256), 0, st, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.n_mic > 0) {
        hipLaunchKernelGGL(k_mic_gather, dim3((a.n_mic + 63) / 64), dim3(64), 0, st, a);
        e = hipGetLastError();
    }
    return e;
}

// One batched step over a table of up to `cap_solves` solves holding at most
// `cap_waves` waves (the grid covers the capacity; n is read on the device).
template <bool NTS>
static hipError_t launch_batch_t(const FusedArgs* table, const int* offsets, const int* n_ptr, int blocks, int rows,
                                 int koff, hipStream_t st) {
    switch (rows) {
        case 2: hipLaunchKernelGGL((k_batch_step<2, NTS>), dim3(blocks), dim3(kBlock), 0, st, table, offsets, n_ptr, koff); break;
        case 4: hipLaunchKernelGGL((k_batch_step<4, NTS>), dim3(blocks), dim3(kBlock), 0, st, table, offsets, n_ptr, koff); break;
        case 8: hipLaunchKernelGGL((k_batch_step<8, NTS>), dim3(blocks), dim3(kBlock), 0, st, table, offsets, n_ptr, koff); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

// Step koff of a chunk. nts: non-temporal stores (a batch whose solves
// together exceed the Infinity Cache), else ordinary stores.
hipError_t launch_batch(const FusedArgs* table, const int* offsets, const int* n_ptr, int cap_waves, int rows,
                        bool nts, int koff, hipStream_t st) {
    int blocks = (cap_waves + kBlockWaves - 1) / kBlockWaves;
    blocks = std::max(8, (blocks + 7) & ~7);
    const hipError_t le = nts ? launch_batch_t<true>(table, offsets, n_ptr, blocks, rows, koff, st)
                              : launch_batch_t<false>(table, offsets, n_ptr, blocks, rows, koff, st);
    if (le != hipSuccess) return le;
    return hipGetLastError();
}

// After the n steps of a chunk: the solves' counters move by n.
hipError_t launch_batch_advance(const FusedArgs* table, const int* n_ptr, int cap_solves, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_advance, dim3((cap_solves + 63) / 64), dim3(64), 0, st, table, n_ptr, n);
    return hipGetLastError();
}

// ---- streaming ceiling ---------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));

// The fused sweep's bytes with none of its stencil: per cell p, vx, vy, vz and
// the flag byte read once from the current set, the four fields written once
// to the other set, four cells per thread in 16-byte accesses. Its rate is the
// practical bandwidth ceiling of the sweep's access mix on the device. The
// written set is the next step's output, so the solve's state is unchanged.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_ceiling(const FusedArgs a, int64_t n4, int tail) {
    const int par = static_cast<int>(*a.step_ptr & 1);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t < n4) {
        const uint32_t f4 = reinterpret_cast<const uint32_t*>(a.flags)[t];
        const float b0 = static_cast<float>(f4 & 0xffu), b1 = static_cast<float>((f4 >> 8) & 0xffu);
        const float b2 = static_cast<float>((f4 >> 16) & 0xffu), b3 = static_cast<float>(f4 >> 24);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const f4v v = reinterpret_cast<const f4v*>(a.f[par][g])[t];
            const f4v o = {v.x + b0, v.y + b1, v.z + b2, v.w + b3};
            f4v* dst = reinterpret_cast<f4v*>(a.f[par ^ 1][g]) + t;
            if (NT) __builtin_nontemporal_store(o, dst);
            else *dst = o;
        }
    } else if (t - n4 < tail) {
        const int64_t c = 4 * n4 + (t - n4);
        const float b = static_cast<float>(a.flags[c]);
#pragma unroll
        for (int g = 0; g < 4; ++g) a.f[par ^ 1][g][c] = a.f[par][g][c] + b;
    }
}

// Enqueue one pass of the streaming kernel, with the sweep's store kind.
hipError_t launch_stream_ceiling(const FusedArgs& a, hipStream_t st) {
    const int64_t cells = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t n4 = cells / 4;
    const int tail = static_cast<int>(cells % 4);
    const int64_t blocks = (n4 + tail + 255) / 256;
    if (cells > kCacheResidentCells)
        hipLaunchKernelGGL(k_stream_ceiling<true>, dim3(blocks), dim3(256), 0, st, a, n4, tail);
    else
        hipLaunchKernelGGL(k_stream_ceiling<false>, dim3(blocks), dim3(256), 0, st, a, n4, tail);
    return hipGetLastError();
}

// The same streaming work in the sweep's own order: lane = k, R rows by
// `planes` planes per wave, the sweep's strip numbering and XCD-aware block
// remap, and no stencil or halo. What the sweep costs beyond this pass is its
// stencil and halo traffic. Grids below 2^30 cells.
template <int R, bool NT>
__global__ __launch_bounds__(kBlock) void k_stream_march(const FusedArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= a.nwaves) return;
    const int seg = w % a.nseg;
    const int t = w / a.nseg;
    const int sj = t % a.nstrip_j, si = t / a.nstrip_j;
    const int k = seg * kWave + lane;
    if (k >= a.nz) return;
    const int i0 = si * a.planes, i1 = min(i0 + a.planes, a.nx);
    const int par = static_cast<int>(*a.step_ptr & 1);
    const GMEM uint8_t* __restrict__ FL = gmem(a.flags);
    for (int i = i0; i < i1; ++i) {
        float v[R][4], b[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = sj * R + r;
            b[r] = 0.f;
            if (j >= a.ny) continue;
            const uint32_t c = lidx<uint32_t>(i, j, k, a.ny, a.nz);
            b[r] = static_cast<float>(ldx(FL, c));
#pragma unroll
            for (int g = 0; g < 4; ++g) v[r][g] = ldx(gmem(static_cast<const float*>(a.f[par][g])), c);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = sj * R + r;
            if (j >= a.ny) continue;
            const uint32_t c = lidx<uint32_t>(i, j, k, a.ny, a.nz);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                if (NT) stx_nt(gmem(a.f[par ^ 1][g]), c, v[r][g] + b[r]);
                else stx(gmem(a.f[par ^ 1][g]), c, v[r][g] + b[r]);
            }
        }
    }
}

template <int R>
static hipError_t launch_march_r(const FusedArgs& a, int blocks, bool nt, hipStream_t st) {
    if (nt) hipLaunchKernelGGL((k_stream_march<R, true>), dim3(blocks), dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((k_stream_march<R, false>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// Enqueue one marching pass with the sweep's rows (2, 4 or 8), planes and
// store kind.
hipError_t launch_stream_march(const FusedArgs& a_in, int rows, int planes, hipStream_t st) {
    FusedArgs a = a_in;
    const int blocks = fused_grid(a, rows, planes, 1);
    const bool nt = static_cast<int64_t>(a.nx) * a.ny * a.nz > kCacheResidentCells;
    switch (rows) {
        case 2: return launch_march_r<2>(a, blocks, nt, st);
        case 4: return launch_march_r<4>(a, blocks, nt, st);
        case 8: return launch_march_r<8>(a, blocks, nt, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_advance(int64_t* step, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, step, n);
    return hipGetLastError();
}

}  // namespace fdtd_amd
