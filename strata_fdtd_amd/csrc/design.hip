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

// (r, kappa) of a voxel and the four Kahan words of its gradients, loaded
// and stored whole
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
// strips: rows per strip of the forward launch (plain and box strips), rows
// per strip of the adjoint's plain strips and of its box strips (the box
// sweep's register budget), planes per strip of the box launches at most.
// Measured at 192^3 with a 128^3 box: the forward is fastest in 2-row strips
// (50 us per step against 60 with 4 rows), the adjoint in 4-row strips with
// 2-row box strips (70.8 us against 77.5 with 2-row strips).
constexpr int kDesignRows = 2;
constexpr int kDesignAdjRows = 4;
constexpr int kDesignAdjBoxRows = 2;
constexpr int kDesignBoxPlanes = 4;

struct DesignArgs {
    // fields: ping-pong sets {p, vx, vy, vz}; step s reads set s & 1
    float* f[2][4];
    const uint8_t* flags;
    const float* dec[3];
    int nx, ny, nz;
    float cv0, cp0;
    // design box [lo, hi) with r = rho/rho0 and kappa = K/K0 (float32)
    int lo[3], hi[3];
    const f2v* rk;                   // device [box][2]: (r, kappa) per voxel, or NULL (air)
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
    // gradient accumulators on the box: Kahan (sum, compensation) of dL/dr,
    // then of dL/dkappa, per voxel
    f4v* acc;
    // pb of every box voxel (the adjoint sweep of step n writes pbb[n & 1]),
    // read by the gradient pass of step n after the next step's sweep
    float* pbb[2];
    // gradient factor -cv / (r_lo + r_hi) of every face of the box, formed once
    // per adjoint call: x faces of cells [lo0, hi0] x [lo1, hi1) x [lo2, hi2)
    // (the face below each cell), likewise y and z; see face_offsets()
    float* face_e;
    // tiling; the strips that meet the design box form the sub-grid
    // [b_seg0, b_seg1) x [b_sj0, b_sj1) x [b_si0, b_si1)
    int planes, nseg, nstrip_j, nwaves, blocks_per_xcd;
    int b_seg0, b_seg1, b_sj0, b_sj1, b_si0, b_si1, b_blocks_per_xcd;
    int rows;  // rows per plain strip; the box strips split them
    // the box launch: planes [b_i0, b_i1) in nb_i strips of box_planes planes
    int box_planes, b_i0, b_i1, nb_i;
};

__device__ __forceinline__ int64_t ebox_cells(const DesignArgs& a) {
    return static_cast<int64_t>(a.hi[0] - a.lo[0] + 2) * (a.hi[1] - a.lo[1] + 2) * (a.hi[2] - a.lo[2] + 2);
}

__device__ __forceinline__ int64_t dbox_cells(const DesignArgs& a) {
    return static_cast<int64_t>(a.hi[0] - a.lo[0]) * (a.hi[1] - a.lo[1]) * (a.hi[2] - a.lo[2]);
}

__device__ __forceinline__ float cvf(const DesignArgs& a, float r_lo, float r_hi) {
    return a.cv0 / ((r_lo + r_hi) * 0.5f);
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

template <typename T>
__device__ __forceinline__ GMEM T* gm(T* p) { return gmem(p); }
template <typename T>
__device__ __forceinline__ const GMEM T* gm(const T* p) { return gmem(p); }

__device__ __forceinline__ float sdown(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float sup(float v) { return __shfl_up(v, 1, kWave); }

__device__ __forceinline__ uint32_t cix(int i, int j, int k, int ny, int nz) { return lidx<uint32_t>(i, j, k, ny, nz); }
template <typename T>
__device__ __forceinline__ T ldo(const GMEM T* base, uint32_t idx) { return ldx(base, idx); }
template <typename T>
__device__ __forceinline__ void sto(GMEM T* base, uint32_t idx, T v) { stx(base, idx, v); }
// A step's field outputs: non-temporal stores for grids beyond the Infinity
// Cache (kCacheResidentCells), ordinary stores below it so that the next step
// finds them on chip.
template <bool NTS, typename T>
__device__ __forceinline__ void stout(GMEM T* base, uint32_t idx, T v) {
    if constexpr (NTS) stx_nt(base, idx, v);
    else stx(base, idx, v);
}

// ---------------------------------------------------------------- box lookups
// Inside the box strips every box quantity a cell needs is loaded at the top
// of its plane, together with the field loads, from an offset clamped into
// the box and then selected (1 or 0 outside), so the loads are unconditional
// and overlap instead of waiting one after another behind branches.

// Position of a coordinate along one axis of the design box [lo, lo + n)
// (off clamped into [0, n)) or of the extended box [lo - 1, lo + n + 1).
struct Ax {
    int off;
    bool in;
};
__device__ __forceinline__ Ax dax(int x, int lo, int n) {
    const int b = x - lo;
    return {min(max(b, 0), n - 1), static_cast<unsigned>(b) < static_cast<unsigned>(n)};
}
__device__ __forceinline__ Ax eax(int x, int lo, int n) {
    const int b = x - lo + 1;
    return {min(max(b, 0), n + 1), static_cast<unsigned>(b) < static_cast<unsigned>(n + 2)};
}

// An absent (r, kappa) array means (1, 1) everywhere: its loads read this.
__device__ __align__(8) float kDesignOne[2] = {1.0f, 1.0f};

struct BoxTab {
    const GMEM f2v* p;
    uint32_t mask;  // offsets are masked to 0 when the array is absent
    bool has;
};
__device__ __forceinline__ BoxTab box_tab(const f2v* p) {
    if (p) return {gm(p), 0xffffffffu, true};
    return {gm(reinterpret_cast<const f2v*>(kDesignOne)), 0u, false};
}
// (r, kappa) of a voxel, (1, 1) outside the box
__device__ __forceinline__ f2v box_val(const BoxTab& t, int off, bool in) {
    const f2v v = ldo(t.p, static_cast<uint32_t>(off) & t.mask);
    const f2v one = {1.0f, 1.0f};
    return (in && t.has) ? v : one;
}

// ---------------------------------------------------------------- forward
// Strips that miss the design box: r = kappa = 1, so every face coefficient
// is cv0 / ((1 + 1) * 0.5) = cv0 exactly and every cell coefficient cp0.
template <int R, bool NTS>
__device__ __forceinline__ void design_forward_strip(const DesignArgs& a, int64_t s, int seg, int j0, int i0, int i1,
                                                     int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const uint32_t sx = static_cast<uint32_t>(ny) * static_cast<uint32_t>(nz);
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
    const double wsrc = a.wave ? a.wave[s] : 0.0;
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    const float cv = a.cv0;
    const uint32_t pmask = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 0);

    bool rvld[R];
    float pc[R], vxm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rvld[r] = j0 + r < ny;
        pc[r] = (rvld[r] && kin) ? ldo(P, cix(i0, j0 + r, k, ny, nz)) : 0.f;
        vxm[r] = 0.f;
        if (i0 > 0 && rvld[r] && kin) {
            const uint32_t c = cix(i0 - 1, j0 + r, k, ny, nz);
            const float v = ldo(VX, c) + cv * (pc[r] - ldo(P, c));
            vxm[r] = (ldo(FL, c) & FDTD_FLAG_OPEN_X) ? v : 0.f;
        }
    }
    float dyr[R];  // the rows' j decays, read once per strip
#pragma unroll
    for (int r = 0; r < R; ++r) dyr[r] = rvld[r] ? a.dec[1][j0 + r] : 1.0f;
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        float pn[R], vxo[R], vyo[R], vzo[R], hr[R], hl_p[R], hl_vz[R];
        uint32_t fl[R], hl_f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pn[r] = vxo[r] = vyo[r] = vzo[r] = hr[r] = hl_p[r] = hl_vz[r] = 0.f;
            fl[r] = hl_f[r] = 0;
            if (!(rvld[r] && kin)) continue;
            const uint32_t c = cix(i, j0 + r, k, ny, nz);
            if (nxt) pn[r] = ldo(P, c + sx);
            vxo[r] = ldo(VX, c);
            vyo[r] = ldo(VY, c);
            vzo[r] = ldo(VZ, c);
            fl[r] = ldo(FL, c);
            if (lane == 0 && k0 > 0) {  // the face below the segment: loads only, used in the row loop
                hl_p[r] = ldo(P, c - 1);
                hl_vz[r] = ldo(VZ, c - 1);
                hl_f[r] = ldo(FL, c - 1);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rvld[r] && lane == kWave - 1 && k + 1 < nz) hr[r] = ldo(P, cix(i, j0 + r, k + 1, ny, nz));
        float p_up_halo = 0.f;
        if (kin && j0 + R < ny) p_up_halo = ldo(P, cix(i, j0 + R, k, ny, nz));
        float vyprev = 0.f;
        if (kin && j0 > 0) {
            const uint32_t c = cix(i, j0 - 1, k, ny, nz);
            const float v = ldo(VY, c) + cv * (pc[0] - ldo(P, c));
            vyprev = (ldo(FL, c) & FDTD_FLAG_OPEN_Y) ? v : 0.f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const uint32_t c = cix(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const float pr = pc[r];
            float vxn = vxo[r];
            if (nxt) {
                const float v = vxo[r] + cv * (pn[r] - pr);
                vxn = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
            float vyn = vyo[r];
            if (j + 1 < ny) {
                const float pu = (r + 1 < R) ? pc[r + 1] : p_up_halo;
                const float v = vyo[r] + cv * (pu - pr);
                vyn = (f & FDTD_FLAG_OPEN_Y) ? v : 0.f;
            }
            float pk1 = sdown(pr);
            if (lane == kWave - 1) pk1 = hr[r];
            float vzn = vzo[r];
            if (k + 1 < nz) {
                const float v = vzo[r] + cv * (pk1 - pr);
                vzn = (f & FDTD_FLAG_OPEN_Z) ? v : 0.f;
            }
            float vzm = sup(vzn);
            if (lane == 0) {
                vzm = 0.f;
                if (k0 > 0 && kin) {
                    const float v = hl_vz[r] + cv * (pr - hl_p[r]);
                    vzm = (hl_f[r] & FDTD_FLAG_OPEN_Z) ? v : 0.f;
                }
            }
            const float div = ((vxn - vxm[r]) + (vyn - vyprev)) + (vzn - vzm);
            float pnew = (f & FDTD_FLAG_AIR) ? pr + a.cp0 * div : 0.f;
            const float dy_j = dyr[r];
            pnew = ((pnew * dx_i) * dy_j) * dz_k;
            if (i == a.src_i && j == a.src_j && k == a.src_k && (f & FDTD_FLAG_AIR))
                pnew = static_cast<float>(static_cast<double>(pnew) + wsrc);
            if (kin) {
                stout<NTS>(Po, c, pnew);
                stout<NTS>(VXo, c, vxn * dx_i);
                stout<NTS>(VYo, c, vyn * dy_j);
                stout<NTS>(VZo, c, vzn * dz_k);
                if (pmask)
                    for (int q = 0; q < a.n_probe; ++q)
                        if (((pmask >> q) & 1u) && a.probe[3 * q] == i && a.probe[3 * q + 1] == j &&
                            a.probe[3 * q + 2] == k)
                            a.trace[s * a.n_probe + q] = pnew;
            }
            vxm[r] = vxn;
            vyprev = vyn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) pc[r] = pn[r];
    }
}

// Strips that meet the design box (with their one-cell halo): per-voxel r and
// kappa, and the forward history of the box. The face coefficient of each
// face is formed from the same two r values in the same order as in the
// per-cell kernel, so the results are the same bits.
template <int R, bool NTS>
__device__ __forceinline__ void design_forward_box_strip(const DesignArgs& a, int64_t s, int seg, int j0, int i0,
                                                         int i1, int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const uint32_t sx = static_cast<uint32_t>(ny) * static_cast<uint32_t>(nz);
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
    GMEM float* hp = hist ? gm(a.hist_p + (s - a.hist0) * ebox_cells(a)) : nullptr;
    GMEM float* hd = hist ? gm(a.hist_div + (s - a.hist0) * dbox_cells(a)) : nullptr;
    const double wsrc = a.wave ? a.wave[s] : 0.0;
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    const uint32_t pmask = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 0);

    const int bx = a.hi[0] - a.lo[0], by = a.hi[1] - a.lo[1], bz = a.hi[2] - a.lo[2];
    const BoxTab RK = box_tab(a.rk);
    // lane halo: lane 0 looks at k - 1, the others at k + 1 (lane 63's is used)
    const int kh = lane == 0 ? k - 1 : k + 1;
    const Ax dk = dax(k, a.lo[2], bz), dkh = dax(kh, a.lo[2], bz), ek = eax(k, a.lo[2], bz);
    Ax dj[R], ej[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dj[r] = dax(j0 + r, a.lo[1], by);
        ej[r] = eax(j0 + r, a.lo[1], by);
    }
    const Ax djd = dax(j0 - 1, a.lo[1], by), dju = dax(j0 + R, a.lo[1], by);
    auto boff = [&](const Ax& x, const Ax& y, const Ax& z) { return (x.off * by + y.off) * bz + z.off; };
    auto RKV = [&](const Ax& x, const Ax& y, const Ax& z) { return box_val(RK, boff(x, y, z), x.in && y.in && z.in); };

    bool rvld[R];
    float pc[R], vxm[R], rcur[R], kcur[R];
    {
        const Ax di0 = dax(i0, a.lo[0], bx), dim = dax(i0 - 1, a.lo[0], bx);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rvld[r] = j0 + r < ny;
            pc[r] = (rvld[r] && kin) ? ldo(P, cix(i0, j0 + r, k, ny, nz)) : 0.f;
            const f2v v0 = RKV(di0, dj[r], dk);
            rcur[r] = v0.x;
            kcur[r] = v0.y;
            const float rm = RKV(dim, dj[r], dk).x;
            vxm[r] = 0.f;
            if (i0 > 0 && rvld[r] && kin) {
                const uint32_t c = cix(i0 - 1, j0 + r, k, ny, nz);
                const float v = ldo(VX, c) + cvf(a, rm, rcur[r]) * (pc[r] - ldo(P, c));
                vxm[r] = (ldo(FL, c) & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
        }
    }
    float dyr[R];  // the rows' j decays, read once per strip
#pragma unroll
    for (int r = 0; r < R; ++r) dyr[r] = rvld[r] ? a.dec[1][j0 + r] : 1.0f;
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        const Ax di = dax(i, a.lo[0], bx), din = dax(i + 1, a.lo[0], bx), ei = eax(i, a.lo[0], bx);
        float pn[R], vxo[R], vyo[R], vzo[R], hr[R], rn[R], kn[R], rkh[R];
        uint32_t fl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pn[r] = vxo[r] = vyo[r] = vzo[r] = hr[r] = 0.f;
            fl[r] = 0;
            const f2v vn = RKV(din, dj[r], dk);
            rn[r] = vn.x;
            kn[r] = vn.y;
            rkh[r] = RKV(di, dj[r], dkh).x;
            if (!(rvld[r] && kin)) continue;
            const uint32_t c = cix(i, j0 + r, k, ny, nz);
            if (nxt) pn[r] = ldo(P, c + sx);
            vxo[r] = ldo(VX, c);
            vyo[r] = ldo(VY, c);
            vzo[r] = ldo(VZ, c);
            fl[r] = ldo(FL, c);
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rvld[r] && lane == kWave - 1 && k + 1 < nz) hr[r] = ldo(P, cix(i, j0 + r, k + 1, ny, nz));
        const float r_up = RKV(di, dju, dk).x, r_dn = RKV(di, djd, dk).x;
        float p_up_halo = 0.f;
        if (kin && j0 + R < ny) p_up_halo = ldo(P, cix(i, j0 + R, k, ny, nz));
        float vyprev = 0.f;
        if (kin && j0 > 0) {
            const uint32_t c = cix(i, j0 - 1, k, ny, nz);
            const float v = ldo(VY, c) + cvf(a, r_dn, rcur[0]) * (pc[0] - ldo(P, c));
            vyprev = (ldo(FL, c) & FDTD_FLAG_OPEN_Y) ? v : 0.f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const uint32_t c = cix(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const float pr = pc[r];
            const float rc = rcur[r];
            const float rc_dn = sdown(rc);  // every lane takes part in the shuffle
            const float r_k1 = (lane == kWave - 1) ? rkh[r] : rc_dn;
            float vxn = vxo[r];
            if (nxt) {
                const float v = vxo[r] + cvf(a, rc, rn[r]) * (pn[r] - pr);
                vxn = (f & FDTD_FLAG_OPEN_X) ? v : 0.f;
            }
            float vyn = vyo[r];
            if (j + 1 < ny) {
                const float pu = (r + 1 < R) ? pc[r + 1] : p_up_halo;
                const float v = vyo[r] + cvf(a, rc, (r + 1 < R) ? rcur[r + 1] : r_up) * (pu - pr);
                vyn = (f & FDTD_FLAG_OPEN_Y) ? v : 0.f;
            }
            float pk1 = sdown(pr);
            if (lane == kWave - 1) pk1 = hr[r];
            float vzn = vzo[r];
            if (k + 1 < nz) {
                const float v = vzo[r] + cvf(a, rc, r_k1) * (pk1 - pr);
                vzn = (f & FDTD_FLAG_OPEN_Z) ? v : 0.f;
            }
            float vzm = sup(vzn);
            if (lane == 0) {
                vzm = 0.f;
                if (k0 > 0 && kin) {
                    // the face below the segment (loaded here: the box strips'
                    // register budget has no room to carry it from the load phase)
                    const uint32_t cl = c - 1;
                    const float v = ldo(VZ, cl) + cvf(a, rkh[r], rc) * (pr - ldo(P, cl));
                    vzm = (ldo(FL, cl) & FDTD_FLAG_OPEN_Z) ? v : 0.f;
                }
            }
            const float div = ((vxn - vxm[r]) + (vyn - vyprev)) + (vzn - vzm);
            float pnew = (f & FDTD_FLAG_AIR) ? pr + (a.cp0 * kcur[r]) * div : 0.f;
            const float dy_j = dyr[r];
            pnew = ((pnew * dx_i) * dy_j) * dz_k;
            if (i == a.src_i && j == a.src_j && k == a.src_k && (f & FDTD_FLAG_AIR))
                pnew = static_cast<float>(static_cast<double>(pnew) + wsrc);
            if (kin) {
                stout<NTS>(Po, c, pnew);
                stout<NTS>(VXo, c, vxn * dx_i);
                stout<NTS>(VYo, c, vyn * dy_j);
                stout<NTS>(VZo, c, vzn * dz_k);
                if (pmask)
                    for (int q = 0; q < a.n_probe; ++q)
                        if (((pmask >> q) & 1u) && a.probe[3 * q] == i && a.probe[3 * q + 1] == j &&
                            a.probe[3 * q + 2] == k)
                            a.trace[s * a.n_probe + q] = pnew;
                if (hist) {
                    if (ei.in && ej[r].in && ek.in)
                        stx_nt(hp, static_cast<uint32_t>((ei.off * (by + 2) + ej[r].off) * (bz + 2) + ek.off), pr);
                    // the history is read back only by the adjoint, much later:
                    // non-temporal, so it does not evict the fields from the caches
                    if (di.in && dj[r].in && dk.in) stx_nt(hd, static_cast<uint32_t>(boff(di, dj[r], dk)), div);
                }
            }
            vxm[r] = vxn;
            vyprev = vyn;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pc[r] = pn[r];
            rcur[r] = rn[r];
            kcur[r] = kn[r];
        }
    }
}

// Strip number w among the strips outside the box sub-grid, numbered in
// (si, sj, seg) order with seg fastest, as in the full grid, so that every XCD
// gets an equal, contiguous share of the work.
__device__ __forceinline__ void plain_strip(const DesignArgs& a, int w, int& seg, int& sj, int& si) {
    const int nseg = a.nseg;
    const int full = nseg * a.nstrip_j;  // strips of a plane range the box misses
    const int bseg = a.b_seg1 - a.b_seg0, bsj = a.b_sj1 - a.b_sj0;
    const int part = full - bseg * bsj;  // strips of a plane range the box meets
    const int nbi = a.b_si1 - a.b_si0;
    if (w < a.b_si0 * full || w >= a.b_si0 * full + nbi * part) {
        if (w >= a.b_si0 * full) w -= nbi * part - nbi * full;
        si = w / full;
        seg = (w % full) % nseg;
        sj = (w % full) / nseg;
        return;
    }
    w -= a.b_si0 * full;
    si = a.b_si0 + w / part;
    int rem = w % part;
    if (rem < a.b_sj0 * nseg) {
        sj = rem / nseg;
        seg = rem % nseg;
        return;
    }
    rem -= a.b_sj0 * nseg;
    const int oseg = nseg - bseg;
    if (rem < bsj * oseg) {
        sj = a.b_sj0 + rem / oseg;
        const int q = rem % oseg;
        seg = q < a.b_seg0 ? q : q + bseg;
        return;
    }
    rem -= bsj * oseg;
    sj = a.b_sj1 + rem / nseg;
    seg = rem % nseg;
}

// The box launch covers the rows [b_sj0, b_sj1) x a.rows and planes
// [b_i0, b_i1) of the box sub-grid in strips of `rows` rows (nb_j of them) and
// box_planes planes (nb_i of them).
__device__ __forceinline__ bool box_strip(const DesignArgs& a, int rows, int w, int& seg, int& j0, int& i0,
                                          int& i1) {
    const int nseg = a.b_seg1 - a.b_seg0, nbj = (a.b_sj1 - a.b_sj0) * (a.rows / rows);
    if (w >= nseg * nbj * a.nb_i) return false;
    const int t = w / nseg;
    seg = a.b_seg0 + w % nseg;
    j0 = a.b_sj0 * a.rows + (t % nbj) * rows;
    if (j0 >= a.ny) return false;  // a split of the last, partial row strip
    i0 = a.b_i0 + (t / nbj) * a.box_planes;
    i1 = min(i0 + a.box_planes, a.b_i1);
    return true;
}

// One launch per step over every strip: the box strips (heavier) first, then
// the others; each XCD gets an equal contiguous share of both.
template <int R, bool NTS>
__global__ __launch_bounds__(kBlock) void k_design_forward(const DesignArgs a, int64_t s) {
    const int xcd = blockIdx.x & 7, bx = blockIdx.x >> 3;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (bx < a.b_blocks_per_xcd) {
        int seg, j0, i0, i1;
        if (box_strip(a, R, (xcd * a.b_blocks_per_xcd + bx) * kBlockWaves + wid, seg, j0, i0, i1))
            design_forward_box_strip<R, NTS>(a, s, seg, j0, i0, i1, lane);
        return;
    }
    const int w = (xcd * a.blocks_per_xcd + bx - a.b_blocks_per_xcd) * kBlockWaves + wid;
    if (w >= a.nwaves) return;
    int seg, sj, si;
    plain_strip(a, w, seg, sj, si);
    const int i0 = si * a.planes;
    design_forward_strip<R, NTS>(a, s, seg, sj * R, i0, min(i0 + a.planes, a.nx), lane);
}

// ---------------------------------------------------------------- adjoint
// pb = dX dY dZ (P + g at probe cells); q = air cp0 kappa pb.
// pm: the strip (with its halo) holds a probe.
// dz: the k decay of the cell (dec[2][k]), kept in a register by the caller.
// The _v forms take the cell's loaded pressure (and flag), so that callers can
// issue those loads together with the plane's other loads.
// dx, dy: the cell's i and j decays (dec[0][i], dec[1][j]).
__device__ __forceinline__ float design_pb_v(const DesignArgs& a, float pv, int i, int j, int k, int64_t n, bool pm,
                                             float dx, float dy, float dz) {
    float v = pv;
    if (pm)
        for (int q = 0; q < a.n_probe; ++q)
            if (a.probe[3 * q] == i && a.probe[3 * q + 1] == j && a.probe[3 * q + 2] == k)
                v += a.g[n * a.n_probe + q];
    const float lam = (dx * dy) * dz;
    return lam * v;
}

__device__ __forceinline__ float design_pb(const DesignArgs& a, const GMEM float* __restrict__ P, int i, int j, int k,
                                           uint32_t c, int64_t n, bool pm, float dz) {
    return design_pb_v(a, ldo(P, c), i, j, k, n, pm, a.dec[0][i], a.dec[1][j], dz);
}

__device__ __forceinline__ float design_q_v(const DesignArgs& a, uint32_t f, float pv, int i, int j, int k, int64_t n,
                                            bool pm, float kap, float dx, float dy, float dz) {
    const float q = a.cp0 * kap * design_pb_v(a, pv, i, j, k, n, pm, dx, dy, dz);
    return (f & FDTD_FLAG_AIR) ? q : 0.f;
}

// Branch-free: the flag and pressure loads are issued together, not one
// after the other behind a test of the flag.
__device__ __forceinline__ float design_q(const DesignArgs& a, const GMEM float* __restrict__ P, int i, int j, int k,
                                          uint32_t c, int64_t n, bool pm, float kap, float dz) {
    return design_q_v(a, ldo(gm(a.flags), c), ldo(P, c), i, j, k, n, pm, kap, a.dec[0][i], a.dec[1][j], dz);
}

// Adjoint step n: (P, V) = lambda_{n+1} in set (n+1)&1 -> lambda_n in set n&1.
// Strips that miss the design box (r = kappa = 1; no gradient).
template <int R, bool NTS>
__device__ __forceinline__ void design_adjoint_strip(const DesignArgs& a, int64_t n, int seg, int j0, int i0, int i1,
                                                     int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const uint32_t sx = static_cast<uint32_t>(ny) * static_cast<uint32_t>(nz);
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
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    // the lane halo's decay: lane 0 at k - 1, the others at k + 1
    const int kh = lane == 0 ? k - 1 : k + 1;
    const float dz_kh = (kh >= 0 && kh < nz) ? a.dec[2][kh] : 1.0f;
    const float cv = a.cv0;
    const bool pm = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 1) != 0;

    bool rvld[R];
    float pbc[R], qc[R], wxm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        rvld[r] = j0 + r < ny;
        pbc[r] = qc[r] = wxm[r] = 0.f;
        if (!(rvld[r] && kin)) continue;
        const int j = j0 + r;
        const uint32_t c = cix(i0, j, k, ny, nz);
        pbc[r] = design_pb(a, P, i0, j, k, c, n, pm, dz_k);
        qc[r] = (ldo(FL, c) & FDTD_FLAG_AIR) ? a.cp0 * pbc[r] : 0.f;
        if (i0 > 0) {
            const uint32_t cl = c - sx;
            const uint32_t fl_m = ldo(FL, cl);
            const float w = a.dec[0][i0 - 1] * ldo(VX, cl) + design_q(a, P, i0 - 1, j, k, cl, n, pm, 1.0f, dz_k) - qc[r];
            wxm[r] = (fl_m & FDTD_FLAG_OPEN_X) ? w : 0.f;
        }
    }
    // decays of the strip's rows and its two halo rows, and per plane of i and i + 1
    float dyr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) dyr[r] = rvld[r] ? a.dec[1][j0 + r] : 1.0f;
    const float dy_up = (j0 + R < ny) ? a.dec[1][j0 + R] : 1.0f;
    const float dy_dn = (j0 > 0) ? a.dec[1][j0 - 1] : 1.0f;
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        const float dx_n = nxt ? a.dec[0][i + 1] : 1.0f;
        // loads only (the halos' arithmetic follows once they are all issued)
        float pbn[R], qn[R], vx[R], vy[R], vz[R], hl_p[R], hl_vz[R], hr_p[R];
        uint32_t fl[R], fln[R], hl_f[R], hr_f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pbn[r] = vx[r] = vy[r] = vz[r] = hl_p[r] = hl_vz[r] = hr_p[r] = 0.f;
            fl[r] = fln[r] = hl_f[r] = hr_f[r] = 0;
            if (!(rvld[r] && kin)) continue;
            const uint32_t c = cix(i, j0 + r, k, ny, nz);
            if (nxt) {
                pbn[r] = ldo(P, c + sx);
                fln[r] = ldo(FL, c + sx);
            }
            vx[r] = ldo(VX, c);
            vy[r] = ldo(VY, c);
            vz[r] = ldo(VZ, c);
            fl[r] = ldo(FL, c);
            if (lane == 0 && k0 > 0) {
                hl_p[r] = ldo(P, c - 1);
                hl_vz[r] = ldo(VZ, c - 1);
                hl_f[r] = ldo(FL, c - 1);
            }
            if (lane == kWave - 1 && k + 1 < nz) {
                hr_p[r] = ldo(P, c + 1);
                hr_f[r] = ldo(FL, c + 1);
            }
        }
        float up_p = 0.f, dn_p = 0.f, dn_vy = 0.f;
        uint32_t up_f = 0, dn_f = 0;
        if (kin && j0 + R < ny) {
            const uint32_t cu = cix(i, j0 + R, k, ny, nz);
            up_p = ldo(P, cu);
            up_f = ldo(FL, cu);
        }
        if (kin && j0 > 0) {
            const uint32_t cl = cix(i, j0 - 1, k, ny, nz);
            dn_p = ldo(P, cl);
            dn_vy = ldo(VY, cl);
            dn_f = ldo(FL, cl);
        }
        float qhr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            qn[r] = qhr[r] = 0.f;
            if (!(rvld[r] && kin)) continue;
            const int j = j0 + r;
            if (nxt) {
                pbn[r] = design_pb_v(a, pbn[r], i + 1, j, k, n, pm, dx_n, dyr[r], dz_k);
                qn[r] = (fln[r] & FDTD_FLAG_AIR) ? a.cp0 * pbn[r] : 0.f;
            }
            qhr[r] = design_q_v(a, hr_f[r], hr_p[r], i, j, k + 1, n, pm, 1.0f, dx_i, dyr[r], dz_kh);
        }
        float q_up_halo = 0.f;
        if (kin && j0 + R < ny) q_up_halo = design_q_v(a, up_f, up_p, i, j0 + R, k, n, pm, 1.0f, dx_i, dy_up, dz_k);
        float wyprev = 0.f;
        if (kin && j0 > 0) {
            const float w = dy_dn * dn_vy + design_q_v(a, dn_f, dn_p, i, j0 - 1, k, n, pm, 1.0f, dx_i, dy_dn, dz_k) - qc[0];
            wyprev = (dn_f & FDTD_FLAG_OPEN_Y) ? w : 0.f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const uint32_t c = cix(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const bool air = f & FDTD_FLAG_AIR;
            float acc = air ? pbc[r] : 0.f;
            float wx, wy, wz;
            const float dy_j = dyr[r];
            if (nxt) {
                wx = (f & FDTD_FLAG_OPEN_X) ? dx_i * vx[r] + qc[r] - qn[r] : 0.f;
                acc -= cv * wx;
            } else {
                wx = dx_i * vx[r] + qc[r];
            }
            if (j + 1 < ny) {
                const float qu = (r + 1 < R) ? qc[r + 1] : q_up_halo;
                wy = (f & FDTD_FLAG_OPEN_Y) ? dy_j * vy[r] + qc[r] - qu : 0.f;
                acc -= cv * wy;
            } else {
                wy = dy_j * vy[r] + qc[r];
            }
            float qk1 = sdown(qc[r]);
            if (lane == kWave - 1) qk1 = qhr[r];
            if (k + 1 < nz) {
                wz = (f & FDTD_FLAG_OPEN_Z) ? dz_k * vz[r] + qc[r] - qk1 : 0.f;
                acc -= cv * wz;
            } else {
                wz = dz_k * vz[r] + qc[r];
            }
            float wzm = sup(wz);
            if (lane == 0) {
                wzm = 0.f;
                if (k0 > 0 && kin) {
                    const float w = dz_kh * hl_vz[r] +
                                    design_q_v(a, hl_f[r], hl_p[r], i, j, k - 1, n, pm, 1.0f, dx_i, dyr[r], dz_kh) -
                                    qc[r];
                    wzm = (hl_f[r] & FDTD_FLAG_OPEN_Z) ? w : 0.f;
                }
            }
            if (i > 0) acc += cv * wxm[r];
            if (j > 0) acc += cv * wyprev;
            if (k > 0) acc += cv * wzm;
            if (kin) {
                stout<NTS>(Po, c, acc);
                stout<NTS>(VXo, c, wx);
                stout<NTS>(VYo, c, wy);
                stout<NTS>(VZo, c, wz);
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

// Strips that meet the design box: per-voxel coefficients. Each face's
// coefficient cv is formed once, where the face is the cell's own (upper)
// face, and passed on to the cell above it: along i in registers from plane
// to plane, along j from row to row, along k from lane to lane. The gradient
// terms are accumulated afterwards by k_design_grad from this step's output
// and the pb of every box voxel, which the strip stores.
template <int R, bool NTS>
__device__ __forceinline__ void design_adjoint_box_strip(const DesignArgs& a, int64_t n, int seg, int j0, int i0,
                                                         int i1, int lane) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int k0 = seg * kWave, k = k0 + lane;
    const bool kin = k < nz;
    const uint32_t sx = static_cast<uint32_t>(ny) * static_cast<uint32_t>(nz);
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
    const float dz_k = kin ? a.dec[2][k] : 1.0f;
    // the lane halo: lane 0 looks at k - 1, the others at k + 1
    const int kh = lane == 0 ? k - 1 : k + 1;
    const float dz_kh = (kh >= 0 && kh < nz) ? a.dec[2][kh] : 1.0f;
    const bool pm = strip_probes(a, i0, i1, j0, j0 + R, k0, k0 + kWave, 1) != 0;

    const int bx = a.hi[0] - a.lo[0], by = a.hi[1] - a.lo[1], bz = a.hi[2] - a.lo[2];
    const BoxTab RK = box_tab(a.rk);
    GMEM float* __restrict__ PBB = gm(a.pbb[n & 1]);
    const Ax dk = dax(k, a.lo[2], bz), dkh = dax(kh, a.lo[2], bz);
    Ax dj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) dj[r] = dax(j0 + r, a.lo[1], by);
    const Ax djd = dax(j0 - 1, a.lo[1], by), dju = dax(j0 + R, a.lo[1], by);
    auto boff = [&](const Ax& x, const Ax& y, const Ax& z) { return (x.off * by + y.off) * bz + z.off; };
    auto RKV = [&](const Ax& x, const Ax& y, const Ax& z) { return box_val(RK, boff(x, y, z), x.in && y.in && z.in); };

    bool rvld[R];
    float pbc[R], qc[R], wxm[R], rcur[R], cvxm[R];
    {
        const Ax di0 = dax(i0, a.lo[0], bx), dim = dax(i0 - 1, a.lo[0], bx);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            rvld[r] = j0 + r < ny;
            pbc[r] = qc[r] = wxm[r] = 0.f;
            const f2v v0 = RKV(di0, dj[r], dk), vm = RKV(dim, dj[r], dk);
            rcur[r] = v0.x;
            cvxm[r] = cvf(a, vm.x, v0.x);
            if (!(rvld[r] && kin)) continue;
            const int j = j0 + r;
            const uint32_t c = cix(i0, j, k, ny, nz);
            pbc[r] = design_pb(a, P, i0, j, k, c, n, pm, dz_k);
            qc[r] = (ldo(FL, c) & FDTD_FLAG_AIR) ? a.cp0 * v0.y * pbc[r] : 0.f;
            if (i0 > 0) {
                const uint32_t cl = c - sx;
                const uint32_t fl_m = ldo(FL, cl);
                const float w = a.dec[0][i0 - 1] * ldo(VX, cl) + design_q(a, P, i0 - 1, j, k, cl, n, pm, vm.y, dz_k) - qc[r];
                wxm[r] = (fl_m & FDTD_FLAG_OPEN_X) ? w : 0.f;
            }
        }
    }
    // decays of the strip's rows and its two halo rows, and per plane of i and i + 1
    float dyr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) dyr[r] = rvld[r] ? a.dec[1][j0 + r] : 1.0f;
    const float dy_up = (j0 + R < ny) ? a.dec[1][j0 + R] : 1.0f;
    const float dy_dn = (j0 > 0) ? a.dec[1][j0 - 1] : 1.0f;
    for (int i = i0; i < i1; ++i) {
        const bool nxt = i + 1 < nx;
        const float dx_i = a.dec[0][i];
        const float dx_n = nxt ? a.dec[0][i + 1] : 1.0f;
        const Ax di = dax(i, a.lo[0], bx), din = dax(i + 1, a.lo[0], bx);
        // loads only (the halos' arithmetic follows once they are all issued)
        float pbn[R], qn[R], vx[R], vy[R], vz[R], rn[R], kn[R], rkh[R], kkh[R], hl_p[R], hl_vz[R], hr_p[R];
        uint32_t fl[R], fln[R], hl_f[R], hr_f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pbn[r] = vx[r] = vy[r] = vz[r] = hl_p[r] = hl_vz[r] = hr_p[r] = 0.f;
            fl[r] = fln[r] = hl_f[r] = hr_f[r] = 0;
            const f2v vn = RKV(din, dj[r], dk), vh = RKV(di, dj[r], dkh);
            rn[r] = vn.x;
            kn[r] = vn.y;
            rkh[r] = vh.x;
            kkh[r] = vh.y;
            if (!(rvld[r] && kin)) continue;
            const uint32_t c = cix(i, j0 + r, k, ny, nz);
            if (nxt) {
                pbn[r] = ldo(P, c + sx);
                fln[r] = ldo(FL, c + sx);
            }
            vx[r] = ldo(VX, c);
            vy[r] = ldo(VY, c);
            vz[r] = ldo(VZ, c);
            fl[r] = ldo(FL, c);
            if (lane == 0 && k0 > 0) {
                hl_p[r] = ldo(P, c - 1);
                hl_vz[r] = ldo(VZ, c - 1);
                hl_f[r] = ldo(FL, c - 1);
            }
            if (lane == kWave - 1 && k + 1 < nz) {
                hr_p[r] = ldo(P, c + 1);
                hr_f[r] = ldo(FL, c + 1);
            }
        }
        const f2v v_up = RKV(di, dju, dk), v_dn = RKV(di, djd, dk);
        float up_p = 0.f, dn_p = 0.f, dn_vy = 0.f;
        uint32_t up_f = 0, dn_f = 0;
        if (kin && j0 + R < ny) {
            const uint32_t cu = cix(i, j0 + R, k, ny, nz);
            up_p = ldo(P, cu);
            up_f = ldo(FL, cu);
        }
        if (kin && j0 > 0) {
            const uint32_t cl = cix(i, j0 - 1, k, ny, nz);
            dn_p = ldo(P, cl);
            dn_vy = ldo(VY, cl);
            dn_f = ldo(FL, cl);
        }
        float qhr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            qn[r] = qhr[r] = 0.f;
            if (!(rvld[r] && kin)) continue;
            const int j = j0 + r;
            if (nxt) {
                pbn[r] = design_pb_v(a, pbn[r], i + 1, j, k, n, pm, dx_n, dyr[r], dz_k);
                qn[r] = (fln[r] & FDTD_FLAG_AIR) ? a.cp0 * kn[r] * pbn[r] : 0.f;
            }
            qhr[r] = design_q_v(a, hr_f[r], hr_p[r], i, j, k + 1, n, pm, kkh[r], dx_i, dyr[r], dz_kh);
        }
        float q_up_halo = 0.f;
        if (kin && j0 + R < ny) q_up_halo = design_q_v(a, up_f, up_p, i, j0 + R, k, n, pm, v_up.y, dx_i, dy_up, dz_k);
        float wyprev = 0.f;
        if (kin && j0 > 0) {
            const float w = dy_dn * dn_vy + design_q_v(a, dn_f, dn_p, i, j0 - 1, k, n, pm, v_dn.y, dx_i, dy_dn, dz_k) -
                            qc[0];
            wyprev = (dn_f & FDTD_FLAG_OPEN_Y) ? w : 0.f;
        }
        // the face below the first row
        float cvyp = cvf(a, v_dn.x, rcur[0]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!rvld[r]) continue;
            const int j = j0 + r;
            const uint32_t c = cix(i, j, k, ny, nz);
            const uint32_t f = fl[r];
            const float rc = rcur[r];
            const bool air = f & FDTD_FLAG_AIR;
            float acc = air ? pbc[r] : 0.f;
            const float rc_dn = sdown(rc);  // every lane takes part in the shuffle
            const float r_k1 = (lane == kWave - 1) ? rkh[r] : rc_dn;
            // own faces (c = lo)
            float wx, wy, wz;
            float cvx = 0.f, cvy = 0.f, cvz = 0.f;
            const float dy_j = dyr[r];
            if (nxt) {
                cvx = cvf(a, rc, rn[r]);
                wx = (f & FDTD_FLAG_OPEN_X) ? dx_i * vx[r] + qc[r] - qn[r] : 0.f;
                acc -= cvx * wx;
            } else {
                wx = dx_i * vx[r] + qc[r];
            }
            if (j + 1 < ny) {
                cvy = cvf(a, rc, (r + 1 < R) ? rcur[r + 1] : v_up.x);
                const float qu = (r + 1 < R) ? qc[r + 1] : q_up_halo;
                wy = (f & FDTD_FLAG_OPEN_Y) ? dy_j * vy[r] + qc[r] - qu : 0.f;
                acc -= cvy * wy;
            } else {
                wy = dy_j * vy[r] + qc[r];
            }
            float qk1 = sdown(qc[r]);
            if (lane == kWave - 1) qk1 = qhr[r];
            if (k + 1 < nz) {
                cvz = cvf(a, rc, r_k1);
                wz = (f & FDTD_FLAG_OPEN_Z) ? dz_k * vz[r] + qc[r] - qk1 : 0.f;
                acc -= cvz * wz;
            } else {
                wz = dz_k * vz[r] + qc[r];
            }
            // faces below (c = hi): the z face from the lane below
            float wzm = sup(wz);
            float cvzm = sup(cvz);
            if (lane == 0) {
                wzm = 0.f;
                cvzm = cvf(a, rkh[r], rc);
                if (k0 > 0 && kin) {
                    const float w =
                        dz_kh * hl_vz[r] + design_q_v(a, hl_f[r], hl_p[r], i, j, k - 1, n, pm, kkh[r], dx_i, dyr[r], dz_kh) -
                        qc[r];
                    wzm = (hl_f[r] & FDTD_FLAG_OPEN_Z) ? w : 0.f;
                }
            }
            if (i > 0) acc += cvxm[r] * wxm[r];
            if (j > 0) acc += cvyp * wyprev;
            if (k > 0) acc += cvzm * wzm;
            if (di.in && dj[r].in && dk.in)
                sto(PBB, static_cast<uint32_t>((di.off * by + dj[r].off) * bz + dk.off), pbc[r]);
            if (kin) {
                stout<NTS>(Po, c, acc);
                stout<NTS>(VXo, c, wx);
                stout<NTS>(VYo, c, wy);
                stout<NTS>(VZo, c, wz);
            }
            wxm[r] = wx;
            cvxm[r] = cvx;
            wyprev = wy;
            cvyp = cvy;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pbc[r] = pbn[r];
            qc[r] = qn[r];
            rcur[r] = rn[r];
        }
    }
}

// Gradient terms of adjoint steps n and, with count 2, n - 1 (in that order),
// one thread per design-box voxel, from each step's output lambda (the w of
// the cell's six faces), its pb (stored by that step's sweep) and the forward
// history of that step:
//   dL/dr     += sum over the six faces of w * (p_hi - p_lo) * (-cv / (r_lo + r_hi))
//   dL/dkappa += air * cp0 * pb * div
// in the per-cell kernel's order (own x, y, z faces, then the faces below),
// with Kahan compensation. Two steps per pass read and write the
// accumulators and the face factors once for both (lambda_n stays in its set
// until the sweep of step n - 2). Every box voxel has all six faces (the box
// keeps one cell to every grid face).
// Layout of face_e: x faces (bx+1) by bz, then y faces bx (by+1) bz, then z
// faces bx by (bz+1); the face of axis d below box cell (bi, bj, bk) (between
// it and its -d neighbour) has that cell's index in the axis' array.
__host__ __device__ inline int64_t face_count(int bx, int by, int bz, int d) {
    return static_cast<int64_t>(bx + (d == 0)) * (by + (d == 1)) * (bz + (d == 2));
}

// The gradient factor of every face of the box: the same expression of the
// same two r values as in the per-cell adjoint kernel.
__global__ __launch_bounds__(256) void k_design_faces(const DesignArgs a) {
    const int bx = a.hi[0] - a.lo[0], by = a.hi[1] - a.lo[1], bz = a.hi[2] - a.lo[2];
    int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    int d = 0;
    while (d < 3 && t >= face_count(bx, by, bz, d)) t -= face_count(bx, by, bz, d++);
    if (d == 3) return;
    const int ey = by + (d == 1), ez = bz + (d == 2);
    const int fk = static_cast<int>(t % ez), fj = static_cast<int>((t / ez) % ey), fi = static_cast<int>(t / (static_cast<int64_t>(ey) * ez));
    const BoxTab RK = box_tab(a.rk);
    auto RV = [&](int ii, int jj, int kk) {
        const bool in = ii >= 0 && ii < bx && jj >= 0 && jj < by && kk >= 0 && kk < bz;
        return box_val(RK, in ? (ii * by + jj) * bz + kk : 0, in).x;
    };
    const float rh = RV(fi, fj, fk);
    const float rl = RV(fi - (d == 0), fj - (d == 1), fk - (d == 2));
    const float cv = cvf(a, rl, rh);
    int64_t off = 0;
    for (int q = 0; q < d; ++q) off += face_count(bx, by, bz, q);
    a.face_e[off + t] = -cv / (rl + rh);
}

__global__ __launch_bounds__(256) void k_design_grad(const DesignArgs a, int64_t n, int count) {
    const int bx = a.hi[0] - a.lo[0], by = a.hi[1] - a.lo[1], bz = a.hi[2] - a.lo[2];
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= static_cast<uint32_t>(bx) * by * bz) return;
    const int bk = static_cast<int>(b % bz), t = static_cast<int>(b / bz);
    const int bj = t % by, bi = t / by;
    const int i = a.lo[0] + bi, j = a.lo[1] + bj, k = a.lo[2] + bk;
    const int ny = a.ny, nz = a.nz;
    const uint32_t sx = static_cast<uint32_t>(ny) * static_cast<uint32_t>(nz);
    const uint32_t c = cix(i, j, k, ny, nz);
    GMEM f4v* __restrict__ ACC = gm(a.acc);
    const int ey = by + 2, ez = bz + 2;
    const uint32_t e = static_cast<uint32_t>(((bi + 1) * ey + (bj + 1)) * ez + (bk + 1));
    // gradient factors of the six faces (below: the face's own cell; above:
    // the +d neighbour's index in the axis' array)
    const GMEM float* __restrict__ EX = gm(a.face_e);
    const GMEM float* __restrict__ EY = EX + face_count(bx, by, bz, 0);
    const GMEM float* __restrict__ EZ = EY + face_count(bx, by, bz, 1);
    const uint32_t fx = static_cast<uint32_t>((bi * by + bj) * bz + bk);
    const uint32_t fy = static_cast<uint32_t>((bi * (by + 1) + bj) * bz + bk);
    const uint32_t fz = static_cast<uint32_t>((bi * by + bj) * (bz + 1) + bk);
    const float exm = ldo(EX, fx), exp_ = ldo(EX, fx + static_cast<uint32_t>(by * bz));
    const float eym = ldo(EY, fy), eyp = ldo(EY, fy + static_cast<uint32_t>(bz));
    const float ezm = ldo(EZ, fz), ezp = ldo(EZ, fz + 1u);
    const bool air = ldo(gm(a.flags), c) & FDTD_FLAG_AIR;
    f4v g = ldo(ACC, b);
    for (int q = 0; q < count; ++q) {
        const int64_t m = n - q;
        const int par = static_cast<int>(m & 1);  // lambda_m lives in set m & 1
        const GMEM float* __restrict__ WX = gm(a.f[par][1]);
        const GMEM float* __restrict__ WY = gm(a.f[par][2]);
        const GMEM float* __restrict__ WZ = gm(a.f[par][3]);
        const GMEM float* __restrict__ hp = gm(a.hist_p + (m - a.hist0) * ebox_cells(a));
        const GMEM float* __restrict__ hd = gm(a.hist_div + (m - a.hist0) * dbox_cells(a));
        const float hpc = ldo(hp, e);
        const float hxp = ldo(hp, e + static_cast<uint32_t>(ey * ez)), hyp = ldo(hp, e + static_cast<uint32_t>(ez)),
                    hzp = ldo(hp, e + 1u);
        const float hxm = ldo(hp, e - static_cast<uint32_t>(ey * ez)), hym = ldo(hp, e - static_cast<uint32_t>(ez)),
                    hzm = ldo(hp, e - 1u);
        const float wx = ldo(WX, c), wy = ldo(WY, c), wz = ldo(WZ, c);
        const float wxm = ldo(WX, c - sx), wym = ldo(WY, c - static_cast<uint32_t>(nz)), wzm = ldo(WZ, c - 1u);
        const float pb = ldo(gm(a.pbb[par]), b);
        const float dv = ldo(hd, b);
        float gr = 0.f;
        gr += wx * (hxp - hpc) * exp_;
        gr += wy * (hyp - hpc) * eyp;
        gr += wz * (hzp - hpc) * ezp;
        gr += wxm * (hpc - hxm) * exm;
        gr += wym * (hpc - hym) * eym;
        gr += wzm * (hpc - hzm) * ezm;
        float y = gr - g.y;
        float s = g.x + y;
        g.y = (s - g.x) - y;
        g.x = s;
        if (air) {
            y = pb * a.cp0 * dv - g.w;
            s = g.z + y;
            g.w = (s - g.z) - y;
            g.z = s;
        }
    }
    sto(ACC, b, g);
}

template <int R, int RB, bool NTS>
__global__ __launch_bounds__(kBlock) void k_design_adjoint(const DesignArgs a, int64_t n) {
    const int xcd = blockIdx.x & 7, bx = blockIdx.x >> 3;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (bx < a.b_blocks_per_xcd) {
        int seg, j0, i0, i1;
        if (box_strip(a, RB, (xcd * a.b_blocks_per_xcd + bx) * kBlockWaves + wid, seg, j0, i0, i1))
            design_adjoint_box_strip<RB, NTS>(a, n, seg, j0, i0, i1, lane);
        return;
    }
    const int w = (xcd * a.blocks_per_xcd + bx - a.b_blocks_per_xcd) * kBlockWaves + wid;
    if (w >= a.nwaves) return;
    int seg, sj, si;
    plain_strip(a, w, seg, sj, si);
    const int i0 = si * a.planes;
    design_adjoint_strip<R, NTS>(a, n, seg, sj * R, i0, min(i0 + a.planes, a.nx), lane);
}

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

// Hardware block b runs on XCD b % 8. Per XCD the launch has
// b_blocks_per_xcd blocks of box strips and blocks_per_xcd of the others;
// returns the number of blocks.
static int design_grid(DesignArgs& a, int planes, int rows, int box_rows) {
    a.rows = rows;
    a.planes = planes;
    a.nseg = (a.nz + kWave - 1) / kWave;
    a.nstrip_j = (a.ny + rows - 1) / rows;
    const int nstrip_i = (a.nx + planes - 1) / planes;
    meeting_strips(a.nz, kWave, a.lo[2], a.hi[2], a.nseg, a.b_seg0, a.b_seg1);
    meeting_strips(a.ny, rows, a.lo[1], a.hi[1], a.nstrip_j, a.b_sj0, a.b_sj1);
    meeting_strips(a.nx, planes, a.lo[0], a.hi[0], nstrip_i, a.b_si0, a.b_si1);
    const int bseg = a.b_seg1 - a.b_seg0, bsj = a.b_sj1 - a.b_sj0, bsi = a.b_si1 - a.b_si0;
    a.nwaves = a.nseg * a.nstrip_j * nstrip_i - bseg * bsj * bsi;
    a.blocks_per_xcd = ((a.nwaves + kBlockWaves - 1) / kBlockWaves + 7) / 8;
    a.box_planes = std::min(planes, kDesignBoxPlanes);
    a.b_i0 = a.b_si0 * planes;
    a.b_i1 = std::min(a.b_si1 * planes, a.nx);
    a.nb_i = bsi ? (a.b_i1 - a.b_i0 + a.box_planes - 1) / a.box_planes : 0;
    const int bw = bseg * bsj * (rows / box_rows) * a.nb_i;
    a.b_blocks_per_xcd = ((bw + kBlockWaves - 1) / kBlockWaves + 7) / 8;
    return 8 * (a.blocks_per_xcd + a.b_blocks_per_xcd);
}

}  // namespace fdtd_amd

using namespace fdtd_amd;

namespace {

int fill_design(DesignArgs& a, float* const* fields, const uint8_t* flags, const float* const* dec, int nx, int ny,
                int nz, float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* rk,
                const int32_t* src_ijk, const double* wave, int n_probe, const int32_t* probe_ijk) {
    if (nx < 2 || ny < 2 || nz < 2 || !fields || !flags || !dec) return -2;
    // 32-bit cell offsets: fewer than 2^30 cells (every byte offset below 4 GiB)
    if (static_cast<int64_t>(nx) * ny * nz >= kMaxCells32) return -3;
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
    // 32-bit box offsets: the 16-byte accumulators of the box below 4 GiB
    if (static_cast<int64_t>(a.hi[0] - a.lo[0] + 2) * (a.hi[1] - a.lo[1] + 2) * (a.hi[2] - a.lo[2] + 2) >= (1ll << 28))
        return -3;
    a.rk = reinterpret_cast<const f2v*>(rk);
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

int64_t fdtd_design_face_count(const int32_t* box_lo, const int32_t* box_hi) {
    const int bx = box_hi[0] - box_lo[0], by = box_hi[1] - box_lo[1], bz = box_hi[2] - box_lo[2];
    return face_count(bx, by, bz, 0) + face_count(bx, by, bz, 1) + face_count(bx, by, bz, 2);
}

int64_t fdtd_design_history_bytes(const int32_t* box_lo, const int32_t* box_hi, int64_t steps) {
    int64_t e = 1, b = 1;
    for (int d = 0; d < 3; ++d) {
        e *= box_hi[d] - box_lo[d] + 2;
        b *= box_hi[d] - box_lo[d];
    }
    return 4 * (e + b) * steps;
}

int fdtd_design_forward(float* const* fields, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* rk,
                        const int32_t* src_ijk, const double* wave, int n_probe, const int32_t* probe_ijk,
                        float* trace, int64_t steps_len, int64_t step0, int n_steps, float* hist_p, float* hist_div,
                        int64_t hist0, int64_t hist_steps, int planes, void* stream) {
    DesignArgs a;
    if (int rc = fill_design(a, fields, flags, dec, nx, ny, nz, cv0, cp0, box_lo, box_hi, rk, src_ijk, wave,
                             n_probe, probe_ijk))
        return rc;
    if (n_steps < 0 || step0 < 0 || planes < 1 || (n_probe > 0 && !trace) || (src_ijk && !wave)) return -2;
    // the trace rows and waveform entries of steps [step0, step0 + n_steps)
    if ((n_probe > 0 || src_ijk) && step0 + n_steps > steps_len) return -4;
    if (hist_p && (!hist_div || hist0 < 0 || hist_steps < 1)) return -2;
    a.trace = trace;
    a.hist_p = hist_p;
    a.hist_div = hist_div;
    a.hist0 = hist0;
    a.hist_steps = hist_p ? hist_steps : 0;
    const int blocks = design_grid(a, planes, kDesignRows, kDesignRows);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool nts = static_cast<int64_t>(nx) * ny * nz > kCacheResidentCells;
    for (int m = 0; m < n_steps; ++m) {
        if (nts)
            hipLaunchKernelGGL((k_design_forward<kDesignRows, true>), dim3(blocks), dim3(kBlock), 0, st, a, step0 + m);
        else
            hipLaunchKernelGGL((k_design_forward<kDesignRows, false>), dim3(blocks), dim3(kBlock), 0, st, a,
                               step0 + m);
        if (hipGetLastError() != hipSuccess) return -5;
    }
    return 0;
}

int fdtd_design_adjoint(float* const* lam, const uint8_t* flags, const float* const* dec, int nx, int ny, int nz,
                        float cv0, float cp0, const int32_t* box_lo, const int32_t* box_hi, const float* rk,
                        int n_probe, const int32_t* probe_ijk, const float* g, int64_t steps_len, int64_t n_lo,
                        int64_t n_hi, const float* hist_p, const float* hist_div, int64_t hist0, int64_t hist_steps,
                        float* acc, float* face_e, float* pb_box, int planes, void* stream) {
    DesignArgs a;
    if (int rc = fill_design(a, lam, flags, dec, nx, ny, nz, cv0, cp0, box_lo, box_hi, rk, nullptr, nullptr, n_probe,
                             probe_ijk))
        return rc;
    if (n_hi < n_lo || n_lo < 0 || planes < 1 || (n_probe > 0 && !g)) return -2;
    if (n_probe > 0 && n_hi > steps_len) return -4;  // rows of g read: [n_lo, n_hi)
    if (!hist_p || !hist_div || !acc || !face_e || !pb_box) return -2;
    if (n_lo < hist0 || n_hi > hist0 + hist_steps) return -2;  // every step needs its history
    a.g = g;
    a.hist_p = const_cast<float*>(hist_p);
    a.hist_div = const_cast<float*>(hist_div);
    a.hist0 = hist0;
    a.hist_steps = hist_steps;
    a.acc = reinterpret_cast<f4v*>(acc);
    a.face_e = face_e;
    const int blocks = design_grid(a, planes, kDesignAdjRows, kDesignAdjBoxRows);
    const int64_t box_cells = static_cast<int64_t>(a.hi[0] - a.lo[0]) * (a.hi[1] - a.lo[1]) * (a.hi[2] - a.lo[2]);
    a.pbb[0] = pb_box;
    a.pbb[1] = pb_box + box_cells;
    const int grad_blocks = static_cast<int>((box_cells + 255) / 256);
    const bool nts = static_cast<int64_t>(nx) * ny * nz > kCacheResidentCells;
    hipStream_t st = static_cast<hipStream_t>(stream);
    {
        const int fbx = a.hi[0] - a.lo[0], fby = a.hi[1] - a.lo[1], fbz = a.hi[2] - a.lo[2];
        const int64_t faces = face_count(fbx, fby, fbz, 0) + face_count(fbx, fby, fbz, 1) + face_count(fbx, fby, fbz, 2);
        hipLaunchKernelGGL(k_design_faces, dim3(static_cast<unsigned>((faces + 255) / 256)), dim3(256), 0, st, a);
        if (hipGetLastError() != hipSuccess) return -5;
    }
    for (int64_t n = n_hi - 1; n >= n_lo; --n) {
        if (nts)
            hipLaunchKernelGGL((k_design_adjoint<kDesignAdjRows, kDesignAdjBoxRows, true>), dim3(blocks), dim3(kBlock),
                               0, st, a, n);
        else
            hipLaunchKernelGGL((k_design_adjoint<kDesignAdjRows, kDesignAdjBoxRows, false>), dim3(blocks),
                               dim3(kBlock), 0, st, a, n);
        // gradient terms two steps per pass: (n + 1, n) after the sweep of n,
        // and a single last step when the count is odd
        const int64_t swept = n_hi - n;
        if (swept % 2 == 0)
            hipLaunchKernelGGL(k_design_grad, dim3(grad_blocks), dim3(256), 0, st, a, n + 1, 2);
        else if (n == n_lo)
            hipLaunchKernelGGL(k_design_grad, dim3(grad_blocks), dim3(256), 0, st, a, n, 1);
        if (hipGetLastError() != hipSuccess) return -5;
    }
    return 0;
}

}  // extern "C"
