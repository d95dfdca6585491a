// One kernel per reference native function, for parity bisection and as the
// per-phase C-ABI entry points. One thread per cell (grid-stride), C order.
// Each kernel states the reference function it replaces.

#include "fdtd_common.hpp"

namespace fdtd_amd {

#define FOR_CELLS(n)                                                                        \
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < (n); \
         c += static_cast<int64_t>(gridDim.x) * blockDim.x)

__device__ __forceinline__ void decompose(int64_t c, int ny, int nz, int& i, int& j, int& k) {
    k = static_cast<int>(c % nz);
    const int64_t t = c / nz;
    j = static_cast<int>(t % ny);
    i = static_cast<int>(t / ny);
}

// fdtd::update_velocity (fdtd_step.cpp:15-82) and the nonuniform form
// (fdtd_step.cpp:235-307): coefficient cv*inv_face formed first.
__global__ void k_velocity(const float* __restrict__ p, float* __restrict__ vx,
                           float* __restrict__ vy, float* __restrict__ vz, int nx, int ny, int nz,
                           float cv, const float* ifx, const float* ify, const float* ifz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz, sx = static_cast<int64_t>(ny) * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        const float pc = p[c];
        if (i < nx - 1) {
            const float cf = ifx ? cv * ifx[i] : cv;
            vx[c] += cf * (p[c + sx] - pc);
        }
        if (j < ny - 1) {
            const float cf = ify ? cv * ify[j] : cv;
            vy[c] += cf * (p[c + nz] - pc);
        }
        if (k < nz - 1) {
            const float cf = ifz ? cv * ifz[k] : cv;
            vz[c] += cf * (p[c + 1] - pc);
        }
    }
}

// fdtd::apply_rigid_boundaries (boundaries.cpp:66-89): the precomputed index
// lists become the OPEN bits of the flag byte.
__global__ void k_rigid(float* vx, float* vy, float* vz, const uint8_t* __restrict__ fl, int nx,
                        int ny, int nz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        const uint32_t f = fl[c];
        if (i < nx - 1 && !(f & FDTD_FLAG_OPEN_X)) vx[c] = 0.f;
        if (j < ny - 1 && !(f & FDTD_FLAG_OPEN_Y)) vy[c] = 0.f;
        if (k < nz - 1 && !(f & FDTD_FLAG_OPEN_Z)) vz[c] = 0.f;
    }
}

// fdtd::update_pressure (fdtd_step.cpp:84-212) and update_pressure_nonuniform
// (:309-439). p is read and written by the same thread only.
__global__ void k_pressure(float* p, const float* __restrict__ vx, const float* __restrict__ vy,
                           const float* __restrict__ vz, const uint8_t* __restrict__ fl, int nx,
                           int ny, int nz, float cp, const float* icx, const float* icy,
                           const float* icz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz, sx = static_cast<int64_t>(ny) * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        if (!(fl[c] & FDTD_FLAG_AIR)) {
            p[c] = 0.f;
            continue;
        }
        const float dvx = vx[c] - (i > 0 ? vx[c - sx] : 0.f);
        const float dvy = vy[c] - (j > 0 ? vy[c - nz] : 0.f);
        const float dvz = vz[c] - (k > 0 ? vz[c - 1] : 0.f);
        if (icx) {
            const float a = icx[i] * dvx, b = icy[j] * dvy, d = icz[k] * dvz;
            p[c] += cp * ((a + b) + d);
        } else {
            p[c] += cp * ((dvx + dvy) + dvz);
        }
    }
}

// fdtd::apply_pml_velocity / apply_pml_pressure (pml.cpp:47-149).
__global__ void k_pml_velocity(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                               const float* dz, int nx, int ny, int nz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        if (dx) vx[c] *= dx[i];
        if (dy) vy[c] *= dy[j];
        if (dz) vz[c] *= dz[k];
    }
}

__global__ void k_pml_pressure(float* p, const float* dx, const float* dy, const float* dz, int nx,
                               int ny, int nz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        float v = p[c];
        if (dx) v *= dx[i];
        if (dy) v *= dy[j];
        if (dz) v *= dz[k];
        p[c] = v;
    }
}

// fdtd::compute_divergence(_nonuniform) (ade.cpp:479-692): d = dvx*ix;
// d += dvy*iy; d += dvz*iz.
__global__ void k_divergence(float* div, const float* __restrict__ vx, const float* __restrict__ vy,
                             const float* __restrict__ vz, int nx, int ny, int nz, float inv_dx,
                             const float* icx, const float* icy, const float* icz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz, sx = static_cast<int64_t>(ny) * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        const float dvx = i > 0 ? vx[c] - vx[c - sx] : vx[c];
        const float dvy = j > 0 ? vy[c] - vy[c - nz] : vy[c];
        const float dvz = k > 0 ? vz[c] - vz[c - 1] : vz[c];
        float d = dvx * (icx ? icx[i] : inv_dx);
        d += dvy * (icy ? icy[j] : inv_dx);
        d += dvz * (icz ? icz[k] : inv_dx);
        div[c] = d;
    }
}

// Flag byte from the geometry mask (precompute_boundary_cells,
// boundaries.cpp:13-64): a face below the last plane is open when both of its
// cells are air.
__global__ void k_build_flags(const uint8_t* __restrict__ air, const uint8_t* __restrict__ mat,
                              uint8_t* __restrict__ fl, int nx, int ny, int nz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz, sx = static_cast<int64_t>(ny) * nz;
    FOR_CELLS(n) {
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        const bool a0 = air[c] != 0;
        uint32_t f = a0 ? FDTD_FLAG_AIR : 0u;
        if (a0 && i < nx - 1 && air[c + sx]) f |= FDTD_FLAG_OPEN_X;
        if (a0 && j < ny - 1 && air[c + nz]) f |= FDTD_FLAG_OPEN_Y;
        if (a0 && k < nz - 1 && air[c + 1]) f |= FDTD_FLAG_OPEN_Z;
        if (mat && mat[c]) f |= FDTD_FLAG_ADE;
        fl[c] = static_cast<uint8_t>(f);
    }
}

// ADE auxiliary update of one pole on the cells of its material
// (ade.cpp:25-222): Debye J = a*J + b*src; Lorentz Jn = a*J + b*Jprev + d*src,
// then Jprev = J, J = Jn. src is p (density poles) or the divergence (modulus).
__global__ void k_ade_update(float* __restrict__ J, float* __restrict__ Jp, const float* __restrict__ src,
                             const uint8_t* __restrict__ mat, int mat_id, int lorentz, float a, float b, float d,
                             int64_t n) {
    FOR_CELLS(n) {
        if (mat[c] != mat_id) continue;
        if (!lorentz) {
            J[c] = a * J[c] + b * src[c];
        } else {
            const float j0 = J[c], j1 = Jp[c];
            const float jn = a * j0 + b * j1 + d * src[c];
            Jp[c] = j0;
            J[c] = jn;
        }
    }
}

// Density-pole velocity correction of one pole (ade.cpp:228-401): on faces
// whose two cells both hold the pole's material, v += coeff * (J[c+e] - J[c]).
__global__ void k_ade_vcorr(float* vx, float* vy, float* vz, const float* __restrict__ J,
                            const uint8_t* __restrict__ mat, int mat_id, float coeff, int nx, int ny, int nz) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz, sx = static_cast<int64_t>(ny) * nz;
    FOR_CELLS(n) {
        if (mat[c] != mat_id) continue;
        int i, j, k;
        decompose(c, ny, nz, i, j, k);
        const float jc = J[c];
        if (i < nx - 1 && mat[c + sx] == mat_id) vx[c] += coeff * (J[c + sx] - jc);
        if (j < ny - 1 && mat[c + nz] == mat_id) vy[c] += coeff * (J[c + nz] - jc);
        if (k < nz - 1 && mat[c + 1] == mat_id) vz[c] += coeff * (J[c + 1] - jc);
    }
}

// Modulus-pole pressure correction of one pole (ade.cpp:403-473):
// p += coeff * J on the cells of its material.
__global__ void k_ade_pcorr(float* p, const float* __restrict__ J, const uint8_t* __restrict__ mat, int mat_id,
                            float coeff, int64_t n) {
    FOR_CELLS(n) {
        if (mat[c] == mat_id) p[c] += coeff * J[c];
    }
}

static inline int grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return static_cast<int>(b < 1 ? 1 : b);
}

hipError_t phase_velocity(const float* p, float* vx, float* vy, float* vz, int nx, int ny, int nz,
                          float cv, const float* ifx, const float* ify, const float* ifz,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_velocity, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, p, vx,
                       vy, vz, nx, ny, nz, cv, ifx, ify, ifz);
    return hipGetLastError();
}

hipError_t phase_rigid(float* vx, float* vy, float* vz, const uint8_t* fl, int nx, int ny, int nz,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_rigid, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, vx, vy,
                       vz, fl, nx, ny, nz);
    return hipGetLastError();
}

hipError_t phase_pressure(float* p, const float* vx, const float* vy, const float* vz,
                          const uint8_t* fl, int nx, int ny, int nz, float cp, const float* icx,
                          const float* icy, const float* icz, hipStream_t st) {
    hipLaunchKernelGGL(k_pressure, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, p, vx,
                       vy, vz, fl, nx, ny, nz, cp, icx, icy, icz);
    return hipGetLastError();
}

hipError_t phase_pml_velocity(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                              const float* dz, int nx, int ny, int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_pml_velocity, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, vx,
                       vy, vz, dx, dy, dz, nx, ny, nz);
    return hipGetLastError();
}

hipError_t phase_pml_pressure(float* p, const float* dx, const float* dy, const float* dz, int nx,
                              int ny, int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_pml_pressure, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, p,
                       dx, dy, dz, nx, ny, nz);
    return hipGetLastError();
}

hipError_t phase_divergence(float* div, const float* vx, const float* vy, const float* vz, int nx,
                            int ny, int nz, float inv_dx, const float* icx, const float* icy,
                            const float* icz, hipStream_t st) {
    hipLaunchKernelGGL(k_divergence, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, div,
                       vx, vy, vz, nx, ny, nz, inv_dx, icx, icy, icz);
    return hipGetLastError();
}

hipError_t phase_build_flags(const uint8_t* air, const uint8_t* mat, uint8_t* fl, int nx, int ny,
                             int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_build_flags, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, air,
                       mat, fl, nx, ny, nz);
    return hipGetLastError();
}

hipError_t phase_ade_update(float* J, float* Jp, const float* src, const uint8_t* mat, int mat_id, int lorentz,
                            float a, float b, float d, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_update, dim3(grid_for(n)), dim3(256), 0, st, J, Jp, src, mat, mat_id, lorentz, a, b, d,
                       n);
    return hipGetLastError();
}

hipError_t phase_ade_vcorr(float* vx, float* vy, float* vz, const float* J, const uint8_t* mat, int mat_id,
                           float coeff, int nx, int ny, int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_vcorr, dim3(grid_for(int64_t(nx) * ny * nz)), dim3(256), 0, st, vx, vy, vz, J, mat,
                       mat_id, coeff, nx, ny, nz);
    return hipGetLastError();
}

hipError_t phase_ade_pcorr(float* p, const float* J, const uint8_t* mat, int mat_id, float coeff, int64_t n,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_ade_pcorr, dim3(grid_for(n)), dim3(256), 0, st, p, J, mat, mat_id, coeff, n);
    return hipGetLastError();
}

}  // namespace fdtd_amd
