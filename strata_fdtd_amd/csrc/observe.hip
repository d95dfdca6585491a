// Observation kernels on the solver state: acoustic energy sums and
// cell-centred velocities. Neither touches the step; both read the field set
// the next step reads (set counter & 1 of the context's device step counter),
// so they can be enqueued between steps without a host sync.
#include <algorithm>

#include "fdtd_common.hpp"

namespace fdtd_amd {

// Energy partial sums, one pair per block: sum over air cells of p^2 and of
// vx^2 + vy^2 + vz^2 (core/solver.py:2689-2706 forms the same sums in float32
// with numpy's pairwise order). Here every square of a float32 is exact in
// float64 and the sums are float64; the grid is fixed and the final kernel
// adds the partials in block order, so a call repeats to the bit.
constexpr int kEnergyBlocks = 1024;
constexpr int kEnergyThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Sum of a block's per-thread pairs; thread 0 returns it.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
    __shared__ double lds[2][kEnergyThreads / kWave];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        lds[0][w] = a;
        lds[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = lds[0][0];
        b = lds[1][0];
        for (int q = 1; q < kEnergyThreads / kWave; ++q) {
            a += lds[0][q];
            b += lds[1][q];
        }
    }
}

__device__ __forceinline__ void energy_cell(uint8_t fl, float p, float vx, float vy, float vz, double& sp,
                                            double& sv) {
    if (fl & FDTD_FLAG_AIR) {
        const double dp = p, dx = vx, dy = vy, dz = vz;
        sp += dp * dp;
        sv += (dx * dx + dy * dy) + dz * dz;
    }
}

// Cells in groups of four (16-byte field loads, one 4-byte flag load), then
// the n % 4 tail. Field buffers 16-byte and flags 4-byte aligned (checked on
// the host).
__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(const FusedArgs a, int64_t n,
                                                                   double* __restrict__ partial) {
    const int par = static_cast<int>(*a.step_ptr & 1);
    const float* __restrict__ P = a.f[par][0];
    const float* __restrict__ VX = a.f[par][1];
    const float* __restrict__ VY = a.f[par][2];
    const float* __restrict__ VZ = a.f[par][3];
    const uint8_t* __restrict__ FL = a.flags;
    const int64_t n4 = n >> 2;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kEnergyThreads;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kEnergyThreads + threadIdx.x;
    double sp = 0.0, sv = 0.0;
    for (int64_t g = t; g < n4; g += stride) {
        const float4 p = reinterpret_cast<const float4*>(P)[g];
        const float4 x = reinterpret_cast<const float4*>(VX)[g];
        const float4 y = reinterpret_cast<const float4*>(VY)[g];
        const float4 z = reinterpret_cast<const float4*>(VZ)[g];
        const uchar4 f = reinterpret_cast<const uchar4*>(FL)[g];
        energy_cell(f.x, p.x, x.x, y.x, z.x, sp, sv);
        energy_cell(f.y, p.y, x.y, y.y, z.y, sp, sv);
        energy_cell(f.z, p.z, x.z, y.z, z.z, sp, sv);
        energy_cell(f.w, p.w, x.w, y.w, z.w, sp, sv);
    }
    for (int64_t c = 4 * n4 + t; c < n; c += stride) energy_cell(FL[c], P[c], VX[c], VY[c], VZ[c], sp, sv);
    block_sum2(sp, sv);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sp;
        partial[2 * blockIdx.x + 1] = sv;
    }
}

// One block: thread t adds partials t, t + 256, ... in order, then the block
// sum; out[0] = sum p^2, out[1] = sum |v|^2.
__global__ __launch_bounds__(kEnergyThreads) void k_energy_final(const double* __restrict__ partial, int nblocks,
                                                                 double* __restrict__ out) {
    double sp = 0.0, sv = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kEnergyThreads) {
        sp += partial[2 * b];
        sv += partial[2 * b + 1];
    }
    block_sum2(sp, sv);
    if (threadIdx.x == 0) {
        out[0] = sp;
        out[1] = sv;
    }
}

int energy_partials() { return kEnergyBlocks; }

hipError_t launch_energy(const FusedArgs& a, double* partial, double* out, hipStream_t st) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t want = (std::max<int64_t>(n >> 2, 1) + kEnergyThreads - 1) / kEnergyThreads;
    const int blocks = static_cast<int>(std::min<int64_t>(want, kEnergyBlocks));
    hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(kEnergyThreads), 0, st, a, n, partial);
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kEnergyThreads), 0, st, partial, blocks, out);
    return hipGetLastError();
}

// Cell-centred velocity along one axis (core/solver.py:2647-2687): out[c] =
// 0.5 * (v[c - e] + v[c]) where the cell's index along the axis is >= 1, and
// 0.5 * v[c] on the first plane. One thread per cell.
__global__ __launch_bounds__(256) void k_centre_velocity(const FusedArgs a, int axis, float* __restrict__ out) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (c >= n) return;
    const int par = static_cast<int>(*a.step_ptr & 1);
    const float* __restrict__ V = a.f[par][1 + axis];
    int64_t e;
    bool first;
    if (axis == 0) {
        e = static_cast<int64_t>(a.ny) * a.nz;
        first = c < e;
    } else if (axis == 1) {
        e = a.nz;
        first = (c / a.nz) % a.ny == 0;
    } else {
        e = 1;
        first = c % a.nz == 0;
    }
    out[c] = first ? 0.5f * V[c] : 0.5f * (V[c - e] + V[c]);
}

hipError_t launch_centre_velocity(const FusedArgs& a, float* ox, float* oy, float* oz, hipStream_t st) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t blocks = (n + 255) / 256;
    float* outs[3] = {ox, oy, oz};
    for (int ax = 0; ax < 3; ++ax) {
        hipLaunchKernelGGL(k_centre_velocity, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, a, ax,
                           outs[ax]);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

}  // namespace fdtd_amd
