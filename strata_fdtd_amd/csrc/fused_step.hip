// Fused FDTD time step for gfx950: velocity update, ADE density correction,
// rigid-wall mask, modulus ADE update, pressure update, PML damping, source
// injection and probe gather in ONE sweep over HBM.
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
// k neighbours inside the segment move across lanes with DPP/ds_bpermute
// shuffles. No LDS, no barriers: waves are independent, so occupancy alone
// hides HBM latency. Ping-pong buffers (read set s&1, write the other) make the
// single sweep race-free.

#include <algorithm>

#include "fdtd_common.hpp"

namespace fdtd_amd {

__device__ __forceinline__ float shfl_down1(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float shfl_up1(float v) { return __shfl_up(v, 1, kWave); }

// Density-pole correction of one face between cells lo and hi of the ADE box.
// Native iteration order: Debye poles, then Lorentz poles (ade.cpp:240-401).
template <bool NU>
__device__ __forceinline__ float face_density_corr(const FusedArgs& a, float v, int m_lo, int m_hi,
                                                   int64_t b_lo, int64_t b_hi, int64_t box_n,
                                                   float inv_face) {
    if (m_lo == 0 || m_lo != m_hi) return v;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = 0; q < a.n_pole; ++q) {
            if (a.pole_target[q] != 0 || a.pole_lorentz[q] != pass || a.pole_mat[q] != m_lo) continue;
            const float* Jq = a.J + q * box_n;
            const float cf = NU ? a.pole_corr[q] * inv_face : a.pole_corr[q];
            v += cf * (Jq[b_hi] - Jq[b_lo]);
        }
    }
    return v;
}

template <int R, bool NU, bool ADE>
__global__ __launch_bounds__(kBlock) void k_fused_step(const FusedArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    // XCD-aware remap: hardware block b runs on XCD b % 8; give each XCD a
    // contiguous range of logical blocks so strips that share halo rows share
    // an L2. gridDim.x is a multiple of 8.
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + (threadIdx.x >> 6);
    int seg, sj, si;
    if (a.sub_mode) {
        const int n0 = a.sub_hi[0] - a.sub_lo[0], n1 = a.sub_hi[1] - a.sub_lo[1],
                  n2 = a.sub_hi[2] - a.sub_lo[2];
        if (w >= n0 * n1 * n2) return;
        seg = a.sub_lo[0] + w % n0;
        const int t = w / n0;
        sj = a.sub_lo[1] + t % n1;
        si = a.sub_lo[2] + t / n1;
    } else {
        if (w >= a.nwaves) return;
        seg = w % a.nseg;
        const int t = w / a.nseg;
        sj = t % a.nstrip_j;
        si = t / a.nstrip_j;
        if (a.skip_sub && seg >= a.sub_lo[0] && seg < a.sub_hi[0] && sj >= a.sub_lo[1] &&
            sj < a.sub_hi[1] && si >= a.sub_lo[2] && si < a.sub_hi[2])
            return;
    }

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave;
    const int k = k0 + lane;
    const bool kv = k < nz;
    const int j0 = sj * R;
    const int i0 = si * a.planes;
    const int i1 = min(i0 + a.planes, nx);
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;

    const int64_t s = *a.step_ptr + a.koff;
    const int par = static_cast<int>(s & 1);
    const float* __restrict__ P = par ? a.f[1][0] : a.f[0][0];
    const float* __restrict__ VX = par ? a.f[1][1] : a.f[0][1];
    const float* __restrict__ VY = par ? a.f[1][2] : a.f[0][2];
    const float* __restrict__ VZ = par ? a.f[1][3] : a.f[0][3];
    float* __restrict__ Po = par ? a.f[0][0] : a.f[1][0];
    float* __restrict__ VXo = par ? a.f[0][1] : a.f[1][1];
    float* __restrict__ VYo = par ? a.f[0][2] : a.f[1][2];
    float* __restrict__ VZo = par ? a.f[0][3] : a.f[1][3];
    const uint8_t* __restrict__ FL = a.flags;
    const float* __restrict__ dX = a.dec[0];
    const float* __restrict__ dY = a.dec[1];
    const float* __restrict__ dZ = a.dec[2];

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
    uint64_t probe_mask = 0;
    for (int q = 0; q < a.n_probe; ++q) {
        const int pi = a.probe_ijk[3 * q], pj = a.probe_ijk[3 * q + 1], pk = a.probe_ijk[3 * q + 2];
        if (pi >= i0 && pi < i1 && pj >= j0 && pj < j0 + R && pk >= k0 && pk < k0 + kWave)
            probe_mask |= 1ull << q;
    }
    const bool sources_ok = s < a.wave_len;
    const bool trace_ok = s < a.trace_len;
    if ((src_mask && !sources_ok) || (probe_mask && !trace_ok)) {
        if (lane == 0) atomicOr(a.health, 2u);
    }

    // ADE box extents (only used when ADE is instantiated)
    int64_t box_n = 0, bsy = 0, bsx = 0;
    bool ade_strip = false;
    if (ADE) {
        const int by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
        bsy = bz;
        bsx = static_cast<int64_t>(by) * bz;
        box_n = bsx * (a.box_hi[0] - a.box_lo[0]);
        ade_strip = (i1 > a.box_lo[0] && i0 - 1 < a.box_hi[0]) &&
                    (j0 + R > a.box_lo[1] && j0 - 1 < a.box_hi[1]) &&
                    (k0 + kWave > a.box_lo[2] && k0 - 1 < a.box_hi[2]);
    }
    auto box_idx = [&](int ii, int jj, int kk) -> int64_t {
        return (ii - a.box_lo[0]) * bsx + (jj - a.box_lo[1]) * bsy + (kk - a.box_lo[2]);
    };
    auto mat_at = [&](int ii, int jj, int kk) -> int {
        return a.material[lin(ii, jj, kk, ny, nz)];
    };

    const float cv = a.cv, cp = a.cp;
    const float ifz_k = (NU && kv && k < nz - 1) ? a.ifc[2][k] : 1.0f;
    const float icz_k = (NU && kv) ? a.icl[2][k] : 1.0f;
    const float cvz = NU ? cv * ifz_k : cv;
    const float dz_k = kv ? dZ[k] : 1.0f;

    float pc[R], pn[R], vxm[R];
    bool rv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rv[r] = (j0 + r) < ny;

    // ---- prologue: p(i0) and the x-face below the strip, vx'(i0-1) -------------
#pragma unroll
    for (int r = 0; r < R; ++r) {
        pc[r] = 0.f;
        vxm[r] = 0.f;
        if (rv[r] && kv) pc[r] = P[lin(i0, j0 + r, k, ny, nz)];
    }
    if (i0 > 0) {
        const int im = i0 - 1;
        const float cvx = NU ? cv * a.ifc[0][im] : cv;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!(rv[r] && kv)) continue;
            const int64_t c = lin(im, j0 + r, k, ny, nz);
            const float pm = P[c], vo = VX[c];
            const uint32_t f = FL[c];
            float v = vo + cvx * (pc[r] - pm);
            if (ADE && ade_strip) {
                const int jr = j0 + r;
                const int m_lo = mat_at(im, jr, k), m_hi = mat_at(i0, jr, k);
                if (m_lo && m_lo == m_hi)
                    v = face_density_corr<NU>(a, v, m_lo, m_hi, box_idx(im, jr, k),
                                              box_idx(i0, jr, k), box_n,
                                              NU ? a.ifc[0][im] : 1.f);
            }
            vxm[r] = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
        }
    }

    // ---- march along i ---------------------------------------------------------
    for (int i = i0; i < i1; ++i) {
        const bool has_next = (i + 1) < nx;
        const float dx_i = dX[i];
        const float cvx = NU ? cv * (has_next ? a.ifc[0][i] : 1.f) : cv;
        const float icx_i = NU ? a.icl[0][i] : 1.f;

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
            const int64_t c = lin(i, j0 + r, k, ny, nz);
            if (kv) {
                if (has_next) pn[r] = P[c + sx];
                vxo[r] = VX[c];
                vyo[r] = VY[c];
                vzo[r] = VZ[c];
                fl[r] = FL[c];
            }
            if (lane == 0 && k0 > 0) {
                hl_p[r] = P[c - 1];
                hl_vz[r] = VZ[c - 1];
                hl_f[r] = FL[c - 1];
            }
            if (lane == kWave - 1 && k + 1 < nz) hr_p[r] = P[c + 1];
        }
        // j halos: p(j0-1), vy(j0-1), flags(j0-1) and p(j0+R)
        float p_jm = 0.f, vy_jm = 0.f, p_jp = 0.f;
        uint32_t f_jm = 0;
        if (kv && j0 > 0) {
            const int64_t c = lin(i, j0 - 1, k, ny, nz);
            p_jm = P[c];
            vy_jm = VY[c];
            f_jm = FL[c];
        }
        if (kv && j0 + R < ny) p_jp = P[lin(i, j0 + R, k, ny, nz)];

        // vy' on the face (j0-1 | j0)
        float vyprev = 0.f;
        if (j0 > 0) {
            const int jm = j0 - 1;
            const float cvy = NU ? cv * a.ifc[1][jm] : cv;
            float v = vy_jm + cvy * (pc[0] - p_jm);
            if (ADE && ade_strip && kv) {
                const int m_lo = mat_at(i, jm, k), m_hi = mat_at(i, j0, k);
                if (m_lo && m_lo == m_hi)
                    v = face_density_corr<NU>(a, v, m_lo, m_hi, box_idx(i, jm, k),
                                              box_idx(i, j0, k), box_n, NU ? a.ifc[1][jm] : 1.f);
            }
            vyprev = (f_jm & FDTD_FLAG_OPEN_Y) ? v : 0.f;
        }

#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rv[r]) continue;  // wave-uniform
            const int jr = j0 + r;
            const int64_t c = lin(i, jr, k, ny, nz);
            const float pr = pc[r];
            const float p_up = (r + 1 < R) ? pc[r + 1] : p_jp;
            const uint32_t f = fl[r];
            const bool row_ade = ADE && ade_strip && i >= a.box_lo[0] - 1 && i < a.box_hi[0] &&
                                 jr >= a.box_lo[1] - 1 && jr < a.box_hi[1];
            int m_c = 0;
            if (ADE && row_ade && kv) m_c = mat_at(i, jr, k);

            // x face (i | i+1)
            float vxn = vxo[r];
            if (has_next) {
                float v = vxo[r] + cvx * (pn[r] - pr);
                if (ADE && row_ade && kv && m_c) {
                    const int m_hi = mat_at(i + 1, jr, k);
                    if (m_hi == m_c)
                        v = face_density_corr<NU>(a, v, m_c, m_hi, box_idx(i, jr, k),
                                                  box_idx(i + 1, jr, k), box_n,
                                                  NU ? a.ifc[0][i] : 1.f);
                }
                vxn = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
            // y face (j | j+1)
            float vyn = vyo[r];
            if (jr + 1 < ny) {
                const float cvy = NU ? cv * a.ifc[1][jr] : cv;
                float v = vyo[r] + cvy * (p_up - pr);
                if (ADE && row_ade && kv && m_c) {
                    const int m_hi = mat_at(i, jr + 1, k);
                    if (m_hi == m_c)
                        v = face_density_corr<NU>(a, v, m_c, m_hi, box_idx(i, jr, k),
                                                  box_idx(i, jr + 1, k), box_n,
                                                  NU ? a.ifc[1][jr] : 1.f);
                }
                vyn = (f & FDTD_FLAG_OPEN_Y) ? v : 0.f;
            }
            // z face (k | k+1): p(k+1) from the next lane, or the right halo
            float pk1 = shfl_down1(pr);
            if (lane == kWave - 1) pk1 = hr_p[r];
            float vzn = vzo[r];
            if (k + 1 < nz) {
                float v = vzo[r] + cvz * (pk1 - pr);
                if (ADE && row_ade && kv && m_c) {
                    const int m_hi = mat_at(i, jr, k + 1);
                    if (m_hi == m_c)
                        v = face_density_corr<NU>(a, v, m_c, m_hi, box_idx(i, jr, k),
                                                  box_idx(i, jr, k + 1), box_n, ifz_k);
                }
                vzn = (f & FDTD_FLAG_OPEN_Z) ? v : 0.f;
            }
            // z face (k-1 | k): previous lane, or the left halo for lane 0
            float vzm = shfl_up1(vzn);
            if (lane == 0) {
                vzm = 0.f;
                if (k0 > 0) {
                    const int km = k0 - 1;
                    const float cz = NU ? cv * a.ifc[2][km] : cv;
                    float v = hl_vz[r] + cz * (pr - hl_p[r]);
                    if (ADE && row_ade && kv) {
                        const int m_lo = mat_at(i, jr, km);
                        if (m_lo && m_lo == m_c)
                            v = face_density_corr<NU>(a, v, m_lo, m_c, box_idx(i, jr, km),
                                                      box_idx(i, jr, k0), box_n,
                                                      NU ? a.ifc[2][km] : 1.f);
                    }
                    vzm = (hl_f[r] & FDTD_FLAG_OPEN_Z) ? v : 0.f;
                }
            }

            // pressure
            const float dvx = vxn - vxm[r];
            const float dvy = vyn - vyprev;
            const float dvz = vzn - vzm;
            float pnew;
            if (NU) {
                const float icy = a.icl[1][jr];
                const float ddx = icx_i * dvx, ddy = icy * dvy, ddz = icz_k * dvz;
                pnew = pr + cp * ((ddx + ddy) + ddz);
            } else {
                pnew = pr + cp * ((dvx + dvy) + dvz);
            }
            if (ADE && row_ade && kv && m_c && i >= a.box_lo[0] && jr >= a.box_lo[1] &&
                k >= a.box_lo[2] && k < a.box_hi[2]) {
                // modulus poles: source is the native divergence of v'
                const float ix = NU ? icx_i : a.ade_inv_dx;
                const float iy = NU ? a.icl[1][jr] : a.ade_inv_dx;
                const float iz = NU ? icz_k : a.ade_inv_dx;
                float dA = dvx * ix;
                dA += dvy * iy;
                dA += dvz * iz;
                const int64_t bc = box_idx(i, jr, k);
#pragma unroll
                for (int pass = 0; pass < 2; ++pass)
                    for (int q = 0; q < a.n_pole; ++q) {
                        if (a.pole_target[q] != 1 || a.pole_lorentz[q] != pass || a.pole_mat[q] != m_c)
                            continue;
                        float* Jq = a.J + q * box_n;
                        if (!pass) {
                            Jq[bc] = a.pole_a[q] * Jq[bc] + a.pole_b[q] * dA;
                        } else {
                            float* Jp = a.Jprev + q * box_n;
                            const float j0v = Jq[bc], j1v = Jp[bc];
                            const float jn = a.pole_a[q] * j0v + a.pole_b[q] * j1v + a.pole_d[q] * dA;
                            Jp[bc] = j0v;
                            Jq[bc] = jn;
                        }
                    }
#pragma unroll
                for (int pass = 0; pass < 2; ++pass)
                    for (int q = 0; q < a.n_pole; ++q) {
                        if (a.pole_target[q] != 1 || a.pole_lorentz[q] != pass || a.pole_mat[q] != m_c)
                            continue;
                        pnew += a.pole_corr[q] * a.J[q * box_n + bc];
                    }
            }
            if (!(f & FDTD_FLAG_AIR)) pnew = 0.f;

            // PML (interior decay is exactly 1.0f, so the multiply is exact there)
            const float dy_j = dY[jr];
            float ps = ((pnew * dx_i) * dy_j) * dz_k;
            if (src_mask && sources_ok && (f & FDTD_FLAG_AIR)) {
                for (int q = 0; q < a.n_src; ++q) {
                    if (!((src_mask >> q) & 1u)) continue;
                    bool on;
                    if (a.src_kind[q] == 0) {
                        on = a.src_i[q] == i && a.src_j[q] == jr && a.src_k[q] == k;
                    } else {
                        const int ax = a.src_i[q], id = a.src_j[q];
                        on = ax == 0 ? id == i : (ax == 1 ? id == jr : id == k);
                    }
                    if (on) ps = static_cast<float>(static_cast<double>(ps) + a.wave[q * a.wave_len + s]);
                }
            }
            if (kv) {
                Po[c] = ps;
                VXo[c] = vxn * dx_i;
                VYo[c] = vyn * dy_j;
                VZo[c] = vzn * dz_k;
            }
            if (probe_mask && trace_ok) {
                for (int q = 0; q < a.n_probe; ++q) {
                    if (!((probe_mask >> q) & 1ull)) continue;
                    if (a.probe_ijk[3 * q] == i && a.probe_ijk[3 * q + 1] == jr &&
                        a.probe_ijk[3 * q + 2] == k) {
                        a.trace[s * a.n_probe + q] = ps;
                        if (!isfinite(ps)) atomicOr(a.health, 1u);
                    }
                }
            }
            vxm[r] = vxn;
            vyprev = vyn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) pc[r] = pn[r];
    }
}

// Density-pole auxiliary update over the ADE box, run before the fused step of
// the same time step (J <- a J + b p, Lorentz with history), ade.cpp:25-166.
__global__ void k_ade_density(const FusedArgs a) {
    const int bx = a.box_hi[0] - a.box_lo[0], by = a.box_hi[1] - a.box_lo[1],
              bz = a.box_hi[2] - a.box_lo[2];
    const int64_t box_n = static_cast<int64_t>(bx) * by * bz;
    const int64_t s = *a.step_ptr + a.koff;
    const float* __restrict__ P = (s & 1) ? a.f[1][0] : a.f[0][0];
    for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; b < box_n;
         b += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int kk = static_cast<int>(b % bz) + a.box_lo[2];
        const int jj = static_cast<int>((b / bz) % by) + a.box_lo[1];
        const int ii = static_cast<int>(b / (static_cast<int64_t>(by) * bz)) + a.box_lo[0];
        const int64_t c = lin(ii, jj, kk, a.ny, a.nz);
        const int m = a.material[c];
        if (!m) continue;
        const float pv = P[c];
        for (int q = 0; q < a.n_pole; ++q) {
            if (a.pole_target[q] != 0 || a.pole_mat[q] != m) continue;
            float* Jq = a.J + q * box_n;
            if (!a.pole_lorentz[q]) {
                Jq[b] = a.pole_a[q] * Jq[b] + a.pole_b[q] * pv;
            } else {
                float* Jp = a.Jprev + q * box_n;
                const float j0v = Jq[b], j1v = Jp[b];
                const float jn = a.pole_a[q] * j0v + a.pole_b[q] * j1v + a.pole_d[q] * pv;
                Jp[b] = j0v;
                Jq[b] = jn;
            }
        }
    }
}

__global__ void k_advance(int64_t* step, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] = step[0] + n;
}

// ---- host-side launch helpers ------------------------------------------------
template <int R, bool NU, bool ADE>
static hipError_t launch_r(const FusedArgs& a, int blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_fused_step<R, NU, ADE>), dim3(blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <int R>
static hipError_t launch_variants(const FusedArgs& a, bool nu, bool ade, int blocks, hipStream_t st) {
    if (nu) return ade ? launch_r<R, true, true>(a, blocks, st) : launch_r<R, true, false>(a, blocks, st);
    return ade ? launch_r<R, false, true>(a, blocks, st) : launch_r<R, false, false>(a, blocks, st);
}

// Fill the tiling fields of `a` for rows R / planes; returns the grid size.
int fused_grid(FusedArgs& a, int rows, int planes) {
    a.planes = planes;
    a.nseg = (a.nz + kWave - 1) / kWave;
    a.nstrip_j = (a.ny + rows - 1) / rows;
    a.nstrip_i = (a.nx + planes - 1) / planes;
    a.nwaves = a.nseg * a.nstrip_j * a.nstrip_i;
    a.sub_mode = 0;
    a.skip_sub = 0;
    int blocks = (a.nwaves + kBlockWaves - 1) / kBlockWaves;
    blocks = (blocks + 7) & ~7;
    a.blocks_per_xcd = blocks / 8;
    return blocks;
}

// With an ADE box the step is two launches over disjoint strips that read the
// same inputs: the plain strips on `st` and the ADE strips on `side`, forked
// after the density-pole update and joined before the step ends, so the small
// ADE launch overlaps the large plain one (also inside a captured graph).
template <int R>
static hipError_t launch_rows(const FusedArgs& a, bool nu, bool ade, int blocks, hipStream_t st,
                              hipStream_t side, hipEvent_t fork, hipEvent_t join) {
    if (!ade) return launch_variants<R>(a, nu, false, blocks, st);
    // Strips that touch the ADE box run the ADE instantiation; every other strip
    // runs the plain one (fewer registers, higher occupancy).
    FusedArgs plain = a;
    FusedArgs mat = a;
    const int rows = R;
    int lo[3] = {a.nseg, a.nstrip_j, a.nstrip_i}, hi[3] = {0, 0, 0};
    for (int s = 0; s < a.nseg; ++s) {
        const int k0 = s * kWave;
        if (k0 + kWave > a.box_lo[2] && k0 - 1 < a.box_hi[2]) { lo[0] = std::min(lo[0], s); hi[0] = s + 1; }
    }
    for (int s = 0; s < a.nstrip_j; ++s) {
        const int j0 = s * rows;
        if (j0 + rows > a.box_lo[1] && j0 - 1 < a.box_hi[1]) { lo[1] = std::min(lo[1], s); hi[1] = s + 1; }
    }
    for (int s = 0; s < a.nstrip_i; ++s) {
        const int i0 = s * a.planes, i1 = std::min(i0 + a.planes, a.nx);
        if (i1 > a.box_lo[0] && i0 - 1 < a.box_hi[0]) { lo[2] = std::min(lo[2], s); hi[2] = s + 1; }
    }
    for (int d = 0; d < 3; ++d) {
        if (hi[d] <= lo[d]) return launch_variants<R>(a, nu, false, blocks, st);
        plain.sub_lo[d] = mat.sub_lo[d] = lo[d];
        plain.sub_hi[d] = mat.sub_hi[d] = hi[d];
    }
    plain.sub_mode = 0;
    plain.skip_sub = 1;
    mat.sub_mode = 1;
    mat.skip_sub = 0;
    const int n_sub = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    int mb = (n_sub + kBlockWaves - 1) / kBlockWaves;
    mb = (mb + 7) & ~7;
    mat.blocks_per_xcd = mb / 8;
    hipError_t e;
    if (side) {
        if ((e = hipEventRecord(fork, st)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(side, fork, 0)) != hipSuccess) return e;
        if ((e = launch_variants<R>(mat, nu, true, mb, side)) != hipSuccess) return e;
        if ((e = hipEventRecord(join, side)) != hipSuccess) return e;
        if ((e = launch_variants<R>(plain, nu, false, blocks, st)) != hipSuccess) return e;
        return hipStreamWaitEvent(st, join, 0);
    }
    if ((e = launch_variants<R>(plain, nu, false, blocks, st)) != hipSuccess) return e;
    return launch_variants<R>(mat, nu, true, mb, st);
}

hipError_t launch_fused(const FusedArgs& a, int rows, bool nu, bool ade, int blocks, hipStream_t st,
                        hipStream_t side, hipEvent_t fork, hipEvent_t join) {
    if (ade) {
        const int64_t bn = static_cast<int64_t>(a.box_hi[0] - a.box_lo[0]) *
                           (a.box_hi[1] - a.box_lo[1]) * (a.box_hi[2] - a.box_lo[2]);
        const int nb = static_cast<int>(std::min<int64_t>((bn + 255) / 256, 4096));
        hipLaunchKernelGGL(k_ade_density, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    switch (rows) {
        case 2: return launch_rows<2>(a, nu, ade, blocks, st, side, fork, join);
        case 4: return launch_rows<4>(a, nu, ade, blocks, st, side, fork, join);
        case 8: return launch_rows<8>(a, nu, ade, blocks, st, side, fork, join);
        case 16: return launch_rows<16>(a, nu, ade, blocks, st, side, fork, join);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_advance(int64_t* step, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, step, n);
    return hipGetLastError();
}

}  // namespace fdtd_amd
