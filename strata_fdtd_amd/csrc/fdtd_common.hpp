// Shared device-side definitions for the MI355X FDTD backend (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fdtd_mi355x.h"

namespace fdtd_amd {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlockWaves = 4;     // waves per workgroup of the fused kernel
constexpr int kBlock = kWave * kBlockWaves;

__host__ __device__ inline int64_t lin(int i, int j, int k, int ny, int nz) {
    return (static_cast<int64_t>(i) * ny + j) * nz + k;
}

// Device (global) memory pointers: saying so lets the compiler use global
// loads with a scalar base even when a pointer comes from memory.
#define GMEM __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GMEM T* gmem(T* p) {
    return (GMEM T*)p;
}
template <typename T>
__device__ __forceinline__ const GMEM T* gmem(const T* p) {
    return (const GMEM T*)p;
}

// Cell offsets of grids below 2^30 cells are 32-bit (uint32_t) and accesses
// with them form the byte offset in 32 bits, so they use the scalar-base +
// 32-bit vector-offset address form: one VGPR per offset, shared by every
// array indexed alike, and no 64-bit address arithmetic per access. int64_t
// offsets index larger grids the ordinary way. ldx / stx / stx_nt pick the
// form from the offset type.
template <typename IX>
__device__ __forceinline__ IX lidx(int i, int j, int k, int ny, int nz) {
    return (static_cast<IX>(i) * static_cast<IX>(ny) + static_cast<IX>(j)) * static_cast<IX>(nz) +
           static_cast<IX>(k);
}
template <typename T>
__device__ __forceinline__ const GMEM T* at32(const GMEM T* base, uint32_t idx) {
    return reinterpret_cast<const GMEM T*>(reinterpret_cast<const GMEM char*>(base) +
                                           idx * static_cast<uint32_t>(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ GMEM T* at32(GMEM T* base, uint32_t idx) {
    return reinterpret_cast<GMEM T*>(reinterpret_cast<GMEM char*>(base) + idx * static_cast<uint32_t>(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T ldx(const GMEM T* base, uint32_t idx) {
    return *at32(base, idx);
}
template <typename T>
__device__ __forceinline__ T ldx(const GMEM T* base, int64_t idx) {
    return base[idx];
}
template <typename T>
__device__ __forceinline__ void stx(GMEM T* base, uint32_t idx, T v) {
    *at32(base, idx) = v;
}
template <typename T>
__device__ __forceinline__ void stx(GMEM T* base, int64_t idx, T v) {
    base[idx] = v;
}
template <typename T>
__device__ __forceinline__ void stx_nt(GMEM T* base, uint32_t idx, T v) {
    __builtin_nontemporal_store(v, at32(base, idx));
}
template <typename T>
__device__ __forceinline__ void stx_nt(GMEM T* base, int64_t idx, T v) {
    __builtin_nontemporal_store(v, base + idx);
}
// grids whose cell offsets fit the 32-bit form
constexpr int64_t kMaxCells32 = int64_t(1) << 30;
// Working sets whose fields and flags (33 B per cell) fit MI355X's 256 MiB
// Infinity Cache stay on chip from one step to the next when the sweep writes
// with ordinary stores (192^3: 154 against 116 Gcell/s with non-temporal
// stores); larger ones are streamed from HBM and written with non-temporal
// stores, which keep L2 for halo re-reads (224^3: 136 against 129).
constexpr int64_t kCacheResidentCells = (int64_t(256) << 20) / 33;

// One ADE pole as the kernels read it (a device table; the list is unbounded).
struct AdePole {
    int32_t mat;                    // material id 1..255
    int32_t target;                 // 0 density, 1 modulus
    int32_t lorentz;                // 0 Debye, 1 Lorentz
    int32_t pad;
    float a, b, d;                  // recurrence coefficients
    float corr;                     // correction coefficient
};

// Pole q of a table that no kernel writes while it runs: read through the
// constant address space, so a wave-uniform q becomes one scalar load (as the
// kernel arguments are) instead of a vector load per lane and a wait.
__device__ __forceinline__ AdePole pole_at(const AdePole* t, int q) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) AdePole*)t)[q];
#else
    return t[q];  // the host pass only parses device functions
#endif
}

// Probes recorded inside the sweep (one wave lane tests one probe); probes
// beyond these are recorded by a small kernel after the step's last pass.
constexpr int kSweepProbes = 64;

// Everything one fused step launch needs. Passed by value (kernarg segment).
struct FusedArgs {
    float* f[2][4];                 // ping-pong sets {p, vx, vy, vz}
    const uint8_t* flags;           // FDTD_FLAG_* per cell
    const float* dec[3];            // PML decay per axis (tables of 1.0f if absent)
    const float* ifc[3];            // nonuniform inverse face spacings (n-1)
    const float* icl[3];            // nonuniform inverse cell spacings (n)
    int nx, ny, nz;
    float cv, cp;                   // coeff_v/coeff_p (uniform) or *_base (nonuniform)

    // launch tiling: wave w -> (segment, j-strip, i-chunk)
    int planes;                     // planes per wave (i extent of a strip)
    int nseg;                       // ceil(nz / 64)
    int nstrip_j;                   // ceil(ny / R)
    int nstrip_i;                   // ceil(nx / planes)
    int nwaves;                     // nseg * nstrip_j * nstrip_i
    int blocks_per_xcd;             // sweep blocks / 8
    int sweep_blocks;               // blocks of the sweep; blocks beyond run
    int dens_blocks;                //   the ADE density update (0: none)

    // time
    const int64_t* step_ptr;        // device counter: index of step koff==0
    int koff;                       // offset of this launch inside a chunk
    int64_t step_limit;             // steps at or past this index are not taken

    // sources
    int n_src;
    int src_kind[FDTD_MAX_SOURCES]; // 0 point, 1 plane
    int src_i[FDTD_MAX_SOURCES];    // point: i ; plane: axis
    int src_j[FDTD_MAX_SOURCES];    // point: j ; plane: index
    int src_k[FDTD_MAX_SOURCES];    // point: k
    const double* wave;             // [n_src][wave_len]
    int64_t wave_len;

    // probes
    int n_probe;                    // all probes (row stride of the trace)
    int n_sweep_probe;              // min(n_probe, kSweepProbes): recorded by the sweep
    const int32_t* probe_ijk;       // device [n_probe][3]
    float* trace;                   // [trace_len][n_probe]
    int64_t trace_len;
    uint32_t* health;               // bit0 non-finite probe, bit1 table overrun

    // microphones: 4 trilinear gathers (p, vx, vy, vz) of 8 corners each
    int n_mic;
    const int32_t* mic_idx;         // device [n_mic][4][8] flat cell indices
    const double* mic_w;            // device [n_mic][4][8] weights
    const int32_t* mic_f64;         // device [n_mic][4]: 1 = float64 sum, 0 = float32 sum
    double* mic_out;                // [mic_len][n_mic][4]
    int64_t mic_len;

    // weighted-mask (membrane) sources, applied after the sweep and the ADE
    // fix-up: per listed cell, its entries in source order, then the probes
    // on that cell are recorded again
    int n_mcell;
    const int32_t* mcell;           // device [n_mcell] flat cell indices, each cell once
    const int32_t* m_estart;        // device [n_mcell + 1] first entry of each cell
    const int32_t* m_esrc;          // device [n_entry] membrane (row of m_wave)
    const int32_t* m_efield;        // device [n_entry] 0 p, 1 vx, 2 vy, 3 vz
    const double* m_eweight;        // device [n_entry] mode-shape weight
    const int32_t* m_pstart;        // device [n_mcell + 1] first probe of each cell
    const int32_t* m_pidx;          // device probe numbers
    const double* m_wave;           // device [n_msrc][m_wave_len] waveform per step
    int64_t m_wave_len;
    int m_single;                   // every listed cell has exactly one entry (entry e of cell e)

    // ADE (materials with poles)
    int n_pole;
    const AdePole* poles;           // device [n_pole], registration order
    const uint8_t* material;        // per cell
    int box_lo[3], box_hi[3];       // ADE box [lo, hi)
    float* J;                       // [n_pole][box cells]
    float* Jprev;                   // [n_pole][box cells] (Lorentz)
    float ade_inv_dx;               // uniform grids: 1/dx of the divergence source
    const int32_t* ade_cells;       // device: box indices of the cells whose material has poles
    int64_t n_ade_cells;            //   (null: the ADE kernels visit every box cell)
};

// The boundary planes of a context (host side of a launch).
struct BoundarySet {
    int n = 0;
    fdtd_boundary_plane planes[FDTD_MAX_BOUNDARY_PLANES];
    float* state = nullptr;
};

}  // namespace fdtd_amd
