// Fast float32 path of the per-voxel design solve and its adjoint (gfx950):
// the forward step and the time-reversed adjoint step as 2.5D i-marching
// sweeps (the layout of the fused solver sweep: a wave owns a 64-wide k
// segment of R rows and marches `planes` planes; k neighbours by lane
// shuffles, j neighbours within the wave or from one halo row, i neighbours
// carried from the previous plane).
//
// Both perform, cell by cell, the operations of the reference-checked kernels
// in adjoint.hip (k_param_step / k_adjoint_step) in the same order, so their
// float32 results are identical to those kernels' (tests compare them bit for
// bit). No counterpart in the reference (scalar rho and c only,
// core/solver.py:1584-1592).
//
// The adjoint needs the forward state only inside the design box: the forward
// sweep records, for step n, p^n on the box grown by one cell on every side
// (the pressure differences across the box cells' faces) and div^n on the box
// (the divergence of the updated velocity), so the adjoint sweep reads no full
// forward state. That history is kept for every step when it fits in HBM
// (no recomputation) or for one checkpoint segment at a time.

#include <algorithm>

#include "fdtd_common.hpp"

namespace fdtd_amd {

constexpr int kDesignMaxProbes = 8;

struct DesignArgs {
    // fields: ping-pong sets {p, vx, vy, vz}; step s reads set s & 1
    float* f[2][4];
    const uint8_t* flags;
    const float* dec[3];
    int nx, ny, nz;
    float cv0, cp0;
    // design box [lo, hi) with r = rho/rho0 and kappa = K/K0 (float32)
    int lo[3], hi[3];
    const float* r;
    const float* kappa;
    // source (one point source; -1 none) and probes
    int src_i, src_j, src_k;
    const double* wave;              // device [steps]
    int n_probe;
    const int32_t* probe;            // device [n_probe][3] (i, j, k)
    float* trace;                    // device [steps][n_probe] (forward)
    const float* g;                  // device [steps][n_probe] (adjoint: dL/dy)
    // history: p on [lo-1, hi+1), div on [lo, hi); slot of step s = s - hist0
    float* hist_p;
    float* hist_div;
    int64_t hist0;
    int64_t hist_steps;
    // gradient accumulators on the box (Kahan sum / compensation)
    float* gr_sum;
    float* gr_comp;
    float* gk_sum;
    float* gk_comp;
    // tiling; the strips that meet the design box form the sub-grid
    // [b_seg0, b_seg1) x [b_sj0, b_sj1) x [b_si0, b_si1)
    int planes, nseg, nstrip_j, nwaves, blocks_per_xcd;
    int b_seg0, b_seg1, b_sj0, b_sj1, b_si0, b_si1, b_blocks_per_xcd;
};

__device__ __forceinline__ bool dbox(const DesignArgs& a, int i, int j, int k) {
    return i >= a.lo[0] && i < a.hi[0] && j >= a.lo[1] && j < a.hi[1] && k >= a.lo[2] && k < a.hi[2];
}

// Box indices fit in 32 bits (checked on the host), which keeps the address
// arithmetic of the many box lookups short.
__device__ __forceinline__ int dbox_lin(const DesignArgs& a, int i, int j, int k) {
    return ((i - a.lo[0]) * (a.hi[1] - a.lo[1]) + (j - a.lo[1])) * (a.hi[2] - a.lo[2]) + (k - a.lo[2]);
}

__device__ __forceinline__ float rv(const DesignArgs& a, int i, int j, int k) {
    return (a.r && dbox(a, i, j, k)) ? a.r[dbox_lin(a, i, j, k)] : 1.0f;
}

__device__ __forceinline__ float kv_(const DesignArgs& a, int i, int j, int k) {
    return (a.kappa && dbox(a, i, j, k)) ? a.kappa[dbox_lin(a, i, j, k)] : 1.0f;
}

// the extended box [lo-1, hi+1) of the pressure history
__device__ __forceinline__ bool ebox(const DesignArgs& a, int i, int j, int k) {
    return i >= a.lo[0] - 1 && i <= a.hi[0] && j >= a.lo[1] - 1 && j <= a.hi[1] && k >= a.lo[2] - 1 &&
           k <= a.hi[2];
}

__device__ __forceinline__ int ebox_lin(const DesignArgs& a, int i, int j, int k) {
    const int ey = a.hi[1] - a.lo[1] + 2, ez = a.hi[2] - a.lo[2] + 2;
    return ((i - a.lo[0] + 1) * ey + (j - a.lo[1] + 1)) * ez + (k - a.lo[2] + 1);
}

__device__ __forceinline__ int64_t ebox_cells(const DesignArgs& a) {
    return static_cast<int64_t>(a.hi[0] - a.lo[0] + 2) * (a.hi[1] - a.lo[1] + 2) * (a.hi[2] - a.lo[2] + 2);
}

__device__ __forceinline__ int64_t dbox_cells(const DesignArgs& a) {
    return static_cast<int64_t>(a.hi[0] - a.lo[0]) * (a.hi[1] - a.lo[1]) * (a.hi[2] - a.lo[2]);
}

__device__ __forceinline__ float cvf(const DesignArgs& a, float r_lo, float r_hi) {
    return a.cv0 / ((r_lo + r_hi) * 0.5f);
}

// Does the region [i0-1, i1] x [j0-1, j1] x [k0-1, k1] (a strip and its
// halo) meet the design box? Outside the box r = kappa = 1, so the face
// coefficient cv0 / ((1 + 1) * 0.5) is exactly cv0 and cp0 * 1 is exactly cp0:
// a strip that misses the box takes those constants without looking anything up.
__device__ __forceinline__ bool strip_meets_box(const DesignArgs& a, int i0, int i1, int j0, int j1, int k0,
                                                int k1) {
    return i0 - 1 < a.hi[0] && i1 >= a.lo[0] && j0 - 1 < a.hi[1] && j1 >= a.lo[1] && k0 - 1 < a.hi[2] &&
           k1 >= a.lo[2];
}

// Probes in [i0-1, i1] x [j0-1, j1] x [k0-1, k1] (halo = 1), or in the strip
// itself [i0, i1) x [j0, j1) x [k0, k1) (halo = 0), as a bit mask.
__device__ __forceinline__ uint32_t strip_probes(const DesignArgs& a, int i0, int i1, int j0, int j1, int k0, int k1,
                                                 int halo) {
    uint32_t m = 0;
    for (int q = 0; q < a.n_probe; ++q) {
        const int pi = a.probe[3 * q], pj = a.probe[3 * q + 1], pk = a.probe[3 * q + 2];
        if (pi >= i0 - halo && pi < i1 + halo && pj >= j0 - halo && pj < j1 + halo && pk >= k0 - halo &&
            pk < k1 + halo)
            m |= 1u << q;
    }
    return m;
}

#define GMEM __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GMEM T* gm(T* p) { return (GMEM T*)p; }
template <typename T>
__device__ __forceinline__ const GMEM T* gm(const T* p) { return (const GMEM T*)p; }

__device__ __forceinline__ float sdown(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float sup(float v) { return __shfl_up(v, 1, kWave); }

// ---------------------------------------------------------------- forward
// TB: the strip (with its halo) meets the design box. Strips that do not
// (most of the grid) run the code with the constant coefficients only.
template <int R, bool TB>
__device__ __forceinline__ void design_forward_strip(const DesignArgs& a, int64_t s, int seg, int sj, int si,
                                                     int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const int j0 = sj * R;
    const int i0 = si * a.planes, i1 = min(i0 + a.planes, nx);
    const int64_t sx = static_cast<int64_t>(ny) * nz;
    const int par = static_cast<int>(s & 1);
    const GMEM float* __restrict__ P = gm(a.f[par][0]);
    const GMEM float* __restrict__ VX = gm(a.f[par][1]);
    const GMEM float* __restrict__ VY = gm(a.f[par][2]);
    const GMEM float* __restrict__ VZ = gm(a.f[par][3]);
    GMEM float* __restrict__ Po = gm(a.f[par ^ 1][0]);
    GMEM float* __restrict__ VXo = gm(a.f[par ^ 1][1]);
    GMEM float* __restrict__ VYo = gm(a.f[par ^ 1][2]);
    GMEM float* __restrict__ VZo = gm(a.f[par ^ 1][3]);
    const GMEM uint8_t* __restrict__ FL = gm(a.flags);
    const bool hist = a.hist_p && s >= a.hist0 && s < a.hist0 + a.hist_steps;
    float* hp = hist ? a.hist_p + (s - a.hist0) * ebox_cells(a) : nullptr;
    float* hd = hist ? a.hist_div + (s - a.hist0) * dbox_cells(a) : nullptr;
    const double wsrc = a.wave ? a.wave[s] : 0.0;
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    constexpr bool tb = TB;
    const uint32_t pmask = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 0);
    auto RV = [&](int ii, int jj, int kk) { return tb ? rv(a, ii, jj, kk) : 1.0f; };
    auto CV = [&](float rl, float rh) { return tb ? cvf(a, rl, rh) : a.cv0; };
    auto KP = [&](int ii, int jj, int kk) { return tb ? kv_(a, ii, jj, kk) : 1.0f; };

    bool rvld[R];
    float pc[R], vxm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rvld[r] = j0 + r < ny;
        pc[r] = (rvld[r] && kin) ? P[lin(i0, j0 + r, k, ny, nz)] : 0.f;
        vxm[r] = 0.f;
        if (i0 > 0 && rvld[r] && kin) {
            const int64_t c = lin(i0 - 1, j0 + r, k, ny, nz);
            const float v = VX[c] + CV(RV(i0 - 1, j0 + r, k), RV(i0, j0 + r, k)) * (pc[r] - P[c]);
            vxm[r] = (FL[c] & FDTD_FLAG_OPEN_X) ? v : 0.f;
        }
    }
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        float pn[R], vxo[R], vyo[R], vzo[R], hr[R];
        uint32_t fl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pn[r] = vxo[r] = vyo[r] = vzo[r] = hr[r] = 0.f;
            fl[r] = 0;
            if (!(rvld[r] && kin)) continue;
            const int64_t c = lin(i, j0 + r, k, ny, nz);
            if (nxt) pn[r] = P[c + sx];
            vxo[r] = VX[c];
            vyo[r] = VY[c];
            vzo[r] = VZ[c];
            fl[r] = FL[c];
        }
        // right k halo (lane 63) and the face below the first row
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rvld[r] && lane == kWave - 1 && k + 1 < nz) hr[r] = P[lin(i, j0 + r, k + 1, ny, nz)];
        float p_up_halo = 0.f;
        if (kin && j0 + R < ny) p_up_halo = P[lin(i, j0 + R, k, ny, nz)];
        float vyprev = 0.f;
        if (kin && j0 > 0) {
            const int64_t c = lin(i, j0 - 1, k, ny, nz);
            const float v = VY[c] + CV(RV(i, j0 - 1, k), RV(i, j0, k)) * (pc[0] - P[c]);
            vyprev = (FL[c] & FDTD_FLAG_OPEN_Y) ? v : 0.f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const int64_t c = lin(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const float pr = pc[r];
            const float rc = RV(i, j, k);
            float vxn = vxo[r];
            if (nxt) {
                const float v = vxo[r] + CV(rc, RV(i + 1, j, k)) * (pn[r] - pr);
                vxn = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
            float vyn = vyo[r];
            if (j + 1 < ny) {
                const float pu = (r + 1 < R) ? pc[r + 1] : p_up_halo;
                const float v = vyo[r] + CV(rc, RV(i, j + 1, k)) * (pu - pr);
                vyn = (f & FDTD_FLAG_OPEN_Y) ? v : 0.f;
            }
            float pk1 = sdown(pr);
            if (lane == kWave - 1) pk1 = hr[r];
            float vzn = vzo[r];
            if (k + 1 < nz) {
                const float v = vzo[r] + CV(rc, RV(i, j, k + 1)) * (pk1 - pr);
                vzn = (f & FDTD_FLAG_OPEN_Z) ? v : 0.f;
            }
            float vzm = sup(vzn);
            if (lane == 0) {
                vzm = 0.f;
                if (k0 > 0 && kin) {
                    const int64_t cl = c - 1;
                    const float v = VZ[cl] + CV(RV(i, j, k - 1), rc) * (pr - P[cl]);
                    vzm = (FL[cl] & FDTD_FLAG_OPEN_Z) ? v : 0.f;
                }
            }
            const float div = ((vxn - vxm[r]) + (vyn - vyprev)) + (vzn - vzm);
            float pnew = (f & FDTD_FLAG_AIR) ? pr + (a.cp0 * KP(i, j, k)) * div : 0.f;
            const float dy_j = a.dec[1][j];
            pnew = ((pnew * dx_i) * dy_j) * dz_k;
            if (i == a.src_i && j == a.src_j && k == a.src_k && (f & FDTD_FLAG_AIR))
                pnew = static_cast<float>(static_cast<double>(pnew) + wsrc);
            if (kin) {
                Po[c] = pnew;
                VXo[c] = vxn * dx_i;
                VYo[c] = vyn * dy_j;
                VZo[c] = vzn * dz_k;
                if (pmask)
                    for (int q = 0; q < a.n_probe; ++q)
                        if (((pmask >> q) & 1u) && a.probe[3 * q] == i && a.probe[3 * q + 1] == j &&
                            a.probe[3 * q + 2] == k)
                            a.trace[s * a.n_probe + q] = pnew;
                if (hist && tb) {
                    if (ebox(a, i, j, k)) hp[ebox_lin(a, i, j, k)] = pr;
                    if (dbox(a, i, j, k)) hd[dbox_lin(a, i, j, k)] = div;
                }
            }
            vxm[r] = vxn;
            vyprev = vyn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) pc[r] = pn[r];
    }
}

// Two launches per step: `box` = 0 runs every strip that misses the design
// box (constant coefficients, the register budget of the plain sweep) and
// skips the others; `box` = 1 runs only the strips that meet it, over the
// sub-grid [b_seg0, b_seg1) x [b_sj0, b_sj1) x [b_si0, b_si1) of strips.
template <int R>
__global__ __launch_bounds__(kBlock) void k_design_forward(const DesignArgs a, int64_t s) {
    const int lane = threadIdx.x & (kWave - 1);
    // XCD-aware: hardware block b runs on XCD b % 8; each XCD gets a contiguous
    // range of strips so neighbouring strips share their halo rows in one L2
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= a.nwaves) return;
    const int seg = w % a.nseg, t = w / a.nseg, sj = t % a.nstrip_j, si = t / a.nstrip_j;
    if (seg >= a.b_seg0 && seg < a.b_seg1 && sj >= a.b_sj0 && sj < a.b_sj1 && si >= a.b_si0 && si < a.b_si1)
        return;
    design_forward_strip<R, false>(a, s, seg, sj, si, lane);
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_design_forward_box(const DesignArgs a, int64_t s) {
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = (blockIdx.x & 7) * a.b_blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nseg = a.b_seg1 - a.b_seg0, nsj = a.b_sj1 - a.b_sj0;
    if (w >= nseg * nsj * (a.b_si1 - a.b_si0)) return;
    const int t = w / nseg;
    design_forward_strip<R, true>(a, s, a.b_seg0 + w % nseg, a.b_sj0 + t % nsj, a.b_si0 + t / nsj, lane);
}

// ---------------------------------------------------------------- adjoint
// pb = dX dY dZ (P + g at probe cells); q = air cp0 kappa pb.
// pm: the strip (with its halo) holds a probe; tb: it meets the design box.
__device__ __forceinline__ float design_pb(const DesignArgs& a, const GMEM float* __restrict__ P, int i, int j, int k,
                                           int64_t c, int64_t n, bool pm) {
    float v = P[c];
    if (pm)
        for (int q = 0; q < a.n_probe; ++q)
            if (a.probe[3 * q] == i && a.probe[3 * q + 1] == j && a.probe[3 * q + 2] == k)
                v += a.g[n * a.n_probe + q];
    const float lam = (a.dec[0][i] * a.dec[1][j]) * a.dec[2][k];
    return lam * v;
}

__device__ __forceinline__ float design_q(const DesignArgs& a, const GMEM float* __restrict__ P, int i, int j, int k,
                                          int64_t c, int64_t n, bool pm, bool tb) {
    if (!(a.flags[c] & FDTD_FLAG_AIR)) return 0.f;
    return a.cp0 * (tb ? kv_(a, i, j, k) : 1.0f) * design_pb(a, P, i, j, k, c, n, pm);
}

__device__ __forceinline__ void kahan(float* sum, float* comp, int64_t b, float x) {
    const float y = x - comp[b];
    const float t = sum[b] + y;
    comp[b] = (t - sum[b]) - y;
    sum[b] = t;
}

// Adjoint step n: (P, V) = lambda_{n+1} in set (n+1)&1 -> lambda_n in set n&1.
template <int R, bool TB>
__device__ __forceinline__ void design_adjoint_strip(const DesignArgs& a, int64_t n, int seg, int sj, int si,
                                                     int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const int j0 = sj * R;
    const int i0 = si * a.planes, i1 = min(i0 + a.planes, nx);
    const int64_t sx = static_cast<int64_t>(ny) * nz;
    const int src = static_cast<int>((n + 1) & 1);
    const GMEM float* __restrict__ P = gm(a.f[src][0]);
    const GMEM float* __restrict__ VX = gm(a.f[src][1]);
    const GMEM float* __restrict__ VY = gm(a.f[src][2]);
    const GMEM float* __restrict__ VZ = gm(a.f[src][3]);
    GMEM float* __restrict__ Po = gm(a.f[src ^ 1][0]);
    GMEM float* __restrict__ VXo = gm(a.f[src ^ 1][1]);
    GMEM float* __restrict__ VYo = gm(a.f[src ^ 1][2]);
    GMEM float* __restrict__ VZo = gm(a.f[src ^ 1][3]);
    const GMEM uint8_t* __restrict__ FL = gm(a.flags);
    const bool grad = a.gr_sum && n >= a.hist0 && n < a.hist0 + a.hist_steps;
    const float* hp = grad ? a.hist_p + (n - a.hist0) * ebox_cells(a) : nullptr;
    const float* hd = grad ? a.hist_div + (n - a.hist0) * dbox_cells(a) : nullptr;
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    constexpr bool tb = TB;
    const bool pm = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 1) != 0;
    auto RV = [&](int ii, int jj, int kk) { return tb ? rv(a, ii, jj, kk) : 1.0f; };
    auto CV = [&](float rl, float rh) { return tb ? cvf(a, rl, rh) : a.cv0; };
    auto KP = [&](int ii, int jj, int kk) { return tb ? kv_(a, ii, jj, kk) : 1.0f; };

    bool rvld[R];
    float pbc[R], qc[R], wxm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rvld[r] = j0 + r < ny;
        pbc[r] = qc[r] = wxm[r] = 0.f;
        if (!(rvld[r] && kin)) continue;
        const int j = j0 + r;
        const int64_t c = lin(i0, j, k, ny, nz);
        pbc[r] = design_pb(a, P, i0, j, k, c, n, pm);
        qc[r] = (FL[c] & FDTD_FLAG_AIR) ? a.cp0 * KP(i0, j, k) * pbc[r] : 0.f;
        if (i0 > 0) {
            const int64_t cl = c - sx;
            wxm[r] = (FL[cl] & FDTD_FLAG_OPEN_X)
                         ? a.dec[0][i0 - 1] * VX[cl] + design_q(a, P, i0 - 1, j, k, cl, n, pm, tb) - qc[r]
                         : 0.f;
        }
    }
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        float pbn[R], qn[R], vx[R], vy[R], vz[R], qhr[R];
        uint32_t fl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pbn[r] = qn[r] = vx[r] = vy[r] = vz[r] = qhr[r] = 0.f;
            fl[r] = 0;
            if (!(rvld[r] && kin)) continue;
            const int j = j0 + r;
            const int64_t c = lin(i, j, k, ny, nz);
            if (nxt) {
                pbn[r] = design_pb(a, P, i + 1, j, k, c + sx, n, pm);
                qn[r] = (FL[c + sx] & FDTD_FLAG_AIR) ? a.cp0 * KP(i + 1, j, k) * pbn[r] : 0.f;
            }
            vx[r] = VX[c];
            vy[r] = VY[c];
            vz[r] = VZ[c];
            fl[r] = FL[c];
            if (lane == kWave - 1 && k + 1 < nz) qhr[r] = design_q(a, P, i, j, k + 1, c + 1, n, pm, tb);
        }
        float q_up_halo = 0.f;
        if (kin && j0 + R < ny) q_up_halo = design_q(a, P, i, j0 + R, k, lin(i, j0 + R, k, ny, nz), n, pm, tb);
        float wyprev = 0.f;
        if (kin && j0 > 0) {
            const int64_t cl = lin(i, j0 - 1, k, ny, nz);
            wyprev = (FL[cl] & FDTD_FLAG_OPEN_Y)
                         ? a.dec[1][j0 - 1] * VY[cl] + design_q(a, P, i, j0 - 1, k, cl, n, pm, tb) - qc[0]
                         : 0.f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const int64_t c = lin(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const float rc = RV(i, j, k);
            const bool air = f & FDTD_FLAG_AIR;
            float acc = air ? pbc[r] : 0.f;
            float gr = 0.f;
            const bool gbox = grad && tb && kin && dbox(a, i, j, k);
            const float hpc = gbox ? hp[ebox_lin(a, i, j, k)] : 0.f;
            // own faces (c = lo)
            float wx, wy, wz;
            const float dy_j = a.dec[1][j];
            if (nxt) {
                const float rh = RV(i + 1, j, k);
                const float cv = CV(rc, rh);
                wx = (f & FDTD_FLAG_OPEN_X) ? dx_i * vx[r] + qc[r] - qn[r] : 0.f;
                acc -= cv * wx;
                if (gbox) gr += wx * (hp[ebox_lin(a, i + 1, j, k)] - hpc) * (-cv / (rc + rh));
            } else {
                wx = dx_i * vx[r] + qc[r];
            }
            if (j + 1 < ny) {
                const float rh = RV(i, j + 1, k);
                const float cv = CV(rc, rh);
                const float qu = (r + 1 < R) ? qc[r + 1] : q_up_halo;
                wy = (f & FDTD_FLAG_OPEN_Y) ? dy_j * vy[r] + qc[r] - qu : 0.f;
                acc -= cv * wy;
                if (gbox) gr += wy * (hp[ebox_lin(a, i, j + 1, k)] - hpc) * (-cv / (rc + rh));
            } else {
                wy = dy_j * vy[r] + qc[r];
            }
            float qk1 = sdown(qc[r]);
            if (lane == kWave - 1) qk1 = qhr[r];
            if (k + 1 < nz) {
                const float rh = RV(i, j, k + 1);
                const float cv = CV(rc, rh);
                wz = (f & FDTD_FLAG_OPEN_Z) ? dz_k * vz[r] + qc[r] - qk1 : 0.f;
                acc -= cv * wz;
                if (gbox) gr += wz * (hp[ebox_lin(a, i, j, k + 1)] - hpc) * (-cv / (rc + rh));
            } else {
                wz = dz_k * vz[r] + qc[r];
            }
            // faces below (c = hi)
            float wzm = sup(wz);
            if (lane == 0) {
                wzm = 0.f;
                if (k0 > 0 && kin) {
                    const int64_t cl = c - 1;
                    wzm = (FL[cl] & FDTD_FLAG_OPEN_Z)
                              ? a.dec[2][k - 1] * VZ[cl] + design_q(a, P, i, j, k - 1, cl, n, pm, tb) - qc[r]
                              : 0.f;
                }
            }
            if (i > 0) {
                const float rl = RV(i - 1, j, k);
                const float cv = CV(rl, rc);
                acc += cv * wxm[r];
                if (gbox) gr += wxm[r] * (hpc - hp[ebox_lin(a, i - 1, j, k)]) * (-cv / (rl + rc));
            }
            if (j > 0) {
                const float rl = RV(i, j - 1, k);
                const float cv = CV(rl, rc);
                acc += cv * wyprev;
                if (gbox) gr += wyprev * (hpc - hp[ebox_lin(a, i, j - 1, k)]) * (-cv / (rl + rc));
            }
            if (k > 0) {
                const float rl = RV(i, j, k - 1);
                const float cv = CV(rl, rc);
                acc += cv * wzm;
                if (gbox) gr += wzm * (hpc - hp[ebox_lin(a, i, j, k - 1)]) * (-cv / (rl + rc));
            }
            if (kin) {
                Po[c] = acc;
                VXo[c] = wx;
                VYo[c] = wy;
                VZo[c] = wz;
                if (gbox) {
                    const int b = dbox_lin(a, i, j, k);
                    kahan(a.gr_sum, a.gr_comp, b, gr);
                    if (air) kahan(a.gk_sum, a.gk_comp, b, pbc[r] * a.cp0 * hd[b]);
                }
            }
            wxm[r] = wx;
            wyprev = wy;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pbc[r] = pbn[r];
            qc[r] = qn[r];
        }
    }
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_design_adjoint(const DesignArgs a, int64_t n) {
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = (blockIdx.x & 7) * a.blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (w >= a.nwaves) return;
    const int seg = w % a.nseg, t = w / a.nseg, sj = t % a.nstrip_j, si = t / a.nstrip_j;
    if (seg >= a.b_seg0 && seg < a.b_seg1 && sj >= a.b_sj0 && sj < a.b_sj1 && si >= a.b_si0 && si < a.b_si1)
        return;
    design_adjoint_strip<R, false>(a, n, seg, sj, si, lane);
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_design_adjoint_box(const DesignArgs a, int64_t n) {
    const int lane = threadIdx.x & (kWave - 1);
    const int lb = (blockIdx.x & 7) * a.b_blocks_per_xcd + (blockIdx.x >> 3);
    const int w = lb * kBlockWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nseg = a.b_seg1 - a.b_seg0, nsj = a.b_sj1 - a.b_sj0;
    if (w >= nseg * nsj * (a.b_si1 - a.b_si0)) return;
    const int t = w / nseg;
    design_adjoint_strip<R, true>(a, n, a.b_seg0 + w % nseg, a.b_sj0 + t % nsj, a.b_si0 + t / nsj, lane);
}

constexpr int kDesignRows = 4;

// Strips [c0, c1) of extent `ext` along an axis of n cells whose span with a
// one-cell halo, [c * ext - 1, min((c + 1) * ext, n)], meets [lo, hi).
static void meeting_strips(int n, int ext, int lo, int hi, int count, int& c0, int& c1) {
    c0 = count;
    c1 = 0;
    for (int c = 0; c < count; ++c) {
        const int a0 = c * ext - 1, a1 = std::min((c + 1) * ext, n);
        if (a0 < hi && a1 >= lo) {
            c0 = std::min(c0, c);
            c1 = c + 1;
        }
    }
    if (c1 <= c0) c0 = c1 = 0;
}

// Returns {blocks of the full-grid launch, blocks of the box launch}.
static void design_grid(DesignArgs& a, int planes, int& blocks, int& box_blocks) {
    a.planes = planes;
    a.nseg = (a.nz + kWave - 1) / kWave;
    a.nstrip_j = (a.ny + kDesignRows - 1) / kDesignRows;
    const int nstrip_i = (a.nx + planes - 1) / planes;
    a.nwaves = a.nseg * a.nstrip_j * nstrip_i;
    blocks = (((a.nwaves + kBlockWaves - 1) / kBlockWaves) + 7) & ~7;
    a.blocks_per_xcd = blocks / 8;
    meeting_strips(a.nz, kWave, a.lo[2], a.hi[2], a.nseg, a.b_seg0, a.b_seg1);
    meeting_strips(a.ny, kDesignRows, a.lo[1], a.hi[1], a.nstrip_j, a.b_sj0, a.b_sj1);
    meeting_strips(a.nx, planes, a.lo[0], a.hi[0], nstrip_i, a.b_si0, a.b_si1);
    const int bw = (a.b_seg1 - a.b_seg0) * (a.b_sj1 - a.b_sj0) * (a.b_si1 - a.b_si0);
    box_blocks = (((bw + kBlockWaves - 1) / kBlockWaves) + 7) & ~7;
    a.b_blocks_per_xcd = box_blocks / 8;
}

}  // namespace fdtd_amd

using namespace fdtd_amd;

namespace {

int fill_design(DesignArgs& a, float* const* fields, const uint8_t* flags, const float* const* dec, int nx, int ny,
                int nz, float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* r,
                const float* kappa, const int32_t* src_ijk, const double* wave, int n_probe,
                const int32_t* probe_ijk) {
    if (nx < 2 || ny < 2 || nz < 2 || !fields || !flags || !dec) return -2;
    if (n_probe < 0 || n_probe > kDesignMaxProbes || (n_probe > 0 && !probe_ijk)) return -2;
    a = DesignArgs{};
    for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 4; ++q) {
            if (!fields[4 * s + q]) return -2;
            a.f[s][q] = fields[4 * s + q];
        }
    a.flags = flags;
    for (int d = 0; d < 3; ++d) {
        if (!dec[d]) return -2;
        a.dec[d] = dec[d];
    }
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.cv0 = cv0;
    a.cp0 = cp0;
    const int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = box_lo ? box_lo[d] : 0;
        a.hi[d] = box_hi ? box_hi[d] : 0;
        if (a.lo[d] < 1 || a.hi[d] > n[d] - 1 || a.hi[d] <= a.lo[d]) return -3;  // box + 1-cell margin inside
    }
    if (static_cast<int64_t>(a.hi[0] - a.lo[0] + 2) * (a.hi[1] - a.lo[1] + 2) * (a.hi[2] - a.lo[2] + 2) >= (1ll << 31))
        return -3;  // box indices are 32-bit
    a.r = r;
    a.kappa = kappa;
    a.src_i = a.src_j = a.src_k = -1;
    if (src_ijk) {
        a.src_i = src_ijk[0];
        a.src_j = src_ijk[1];
        a.src_k = src_ijk[2];
    }
    a.wave = wave;
    a.n_probe = n_probe;
    a.probe = probe_ijk;  // device; only compared with cell coordinates
    return 0;
}

}  // namespace

extern "C" {

int64_t fdtd_design_history_bytes(const int32_t* box_lo, const int32_t* box_hi, int64_t steps) {
    int64_t e = 1, b = 1;
    for (int d = 0; d < 3; ++d) {
        e *= box_hi[d] - box_lo[d] + 2;
        b *= box_hi[d] - box_lo[d];
    }
    return 4 * (e + b) * steps;
}

int fdtd_design_forward(float* const* fields, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* r,
                        const float* kappa, const int32_t* src_ijk, const double* wave, int n_probe,
                        const int32_t* probe_ijk, float* trace, int64_t step0, int n_steps, float* hist_p,
                        float* hist_div, int64_t hist0, int64_t hist_steps, int planes, void* stream) {
    DesignArgs a;
    if (int rc = fill_design(a, fields, flags, dec, nx, ny, nz, cv0, cp0, box_lo, box_hi, r, kappa, src_ijk, wave,
                             n_probe, probe_ijk))
        return rc;
    if (n_steps < 0 || step0 < 0 || planes < 1 || (n_probe > 0 && !trace) || (src_ijk && !wave)) return -2;
    if (hist_p && (!hist_div || hist0 < 0 || hist_steps < 1)) return -2;
    a.trace = trace;
    a.hist_p = hist_p;
    a.hist_div = hist_div;
    a.hist0 = hist0;
    a.hist_steps = hist_p ? hist_steps : 0;
    int blocks = 0, box_blocks = 0;
    design_grid(a, planes, blocks, box_blocks);
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (int m = 0; m < n_steps; ++m) {
        hipLaunchKernelGGL(k_design_forward<kDesignRows>, dim3(blocks), dim3(kBlock), 0, st, a, step0 + m);
        if (box_blocks)
            hipLaunchKernelGGL(k_design_forward_box<kDesignRows>, dim3(box_blocks), dim3(kBlock), 0, st, a,
                               step0 + m);
        if (hipGetLastError() != hipSuccess) return -5;
    }
    return 0;
}

int fdtd_design_adjoint(float* const* lam, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* r,
                        const float* kappa, int n_probe, const int32_t* probe_ijk, const float* g, int64_t n_lo,
                        int64_t n_hi, const float* hist_p, const float* hist_div, int64_t hist0, int64_t hist_steps,
                        float* gr_sum, float* gr_comp, float* gk_sum, float* gk_comp, int planes, void* stream) {
    DesignArgs a;
    if (int rc = fill_design(a, lam, flags, dec, nx, ny, nz, cv0, cp0, box_lo, box_hi, r, kappa, nullptr, nullptr,
                             n_probe, probe_ijk))
        return rc;
    if (n_hi < n_lo || n_lo < 0 || planes < 1 || (n_probe > 0 && !g)) return -2;
    if (!hist_p || !hist_div || !gr_sum || !gr_comp || !gk_sum || !gk_comp) return -2;
    if (n_lo < hist0 || n_hi > hist0 + hist_steps) return -2;  // every step needs its history
    a.g = g;
    a.hist_p = const_cast<float*>(hist_p);
    a.hist_div = const_cast<float*>(hist_div);
    a.hist0 = hist0;
    a.hist_steps = hist_steps;
    a.gr_sum = gr_sum;
    a.gr_comp = gr_comp;
    a.gk_sum = gk_sum;
    a.gk_comp = gk_comp;
    int blocks = 0, box_blocks = 0;
    design_grid(a, planes, blocks, box_blocks);
    hipStream_t st = static_cast<hipStream_t>(stream);
    for (int64_t n = n_hi - 1; n >= n_lo; --n) {
        hipLaunchKernelGGL(k_design_adjoint<kDesignRows>, dim3(blocks), dim3(kBlock), 0, st, a, n);
        if (box_blocks)
            hipLaunchKernelGGL(k_design_adjoint_box<kDesignRows>, dim3(box_blocks), dim3(kBlock), 0, st, a, n);
        if (hipGetLastError() != hipSuccess) return -5;
    }
    return 0;
}

}  // extern "C"
