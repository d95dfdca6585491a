// Device voxeliser: CSG trees of boxes, spheres and axis-aligned cylinders and cones,
// translated and scaled, evaluated in float64 at every cell centre, giving the air mask (sdf <= 0)
// and the material ids of MaterializedGeometry, bit-identical to the
// reference's numpy evaluation.
//
// Reference formulas (geometry/sdf.py): Box :274-294, Sphere :320-333,
// Cylinder :382-425, Cone :436-543, Union/Intersection/Difference :566-722,
// Translate :1199-1237, Scale :1239-1301, voxelize :99-127;
// material ids geometry/material_assignment.py:133-168 (later regions
// override earlier ones, only inside air).
//
// Exactness. Every float64 value a test depends on is formed with the same
// IEEE operations in the same order as numpy forms it: np.linalg.norm over
// axis 1 is sqrt((x0*x0 + x1*x1) + x2*x2); `** 2` is x*x; the cylinder's and
// cone's np.dot with an axis-aligned unit vector is exactly +-v_axis (every
// other product is a signed zero); sqrt is correctly rounded in both. The file is
// built with -ffp-contract=off, so no multiply-add is fused. CSG nodes
// combine the per-primitive tests as booleans: min(a, b) <= 0 iff a <= 0 or
// b <= 0; max(a, b) <= 0 iff both; max(b, -s) <= 0 iff b <= 0 and not s < 0
// (strict forms dually), so no intermediate float64 value of a CSG node is
// needed. A box is inside iff every |x - c| - size/2 is <= 0 (< 0 strict).
// A translation evaluates its subtree at (x - ox, y - oy, z - oz) and a scale
// at (x / sx, y / sy, z / sz), the coordinates numpy forms (`points - offset`,
// `points / scale`); a scale multiplies its subtree's distance by a positive
// factor, which keeps the sign (the factors are limited to [1e-6, 1e6], so
// the product of a nonzero distance never underflows to zero).
//
// Program format (postfix, validated on the host before launch). Each
// instruction is two int32: {op, arg}.
//   VOX_BOX  arg = offset of {cx, cy, cz, hx, hy, hz} in params   push test
//   VOX_SPH  arg = offset of {cx, cy, cz, r}                      push test
//   VOX_CYL  arg = offset of {p1x, p1y, p1z, nx, ny, nz, len, r, axis}
//   VOX_CONE arg = offset of {p1x, p1y, p1z, nx, ny, nz, len, r1, r2, axis}
//   the strict variants (+ VOX_STRICT) test sdf < 0 instead of sdf <= 0
//   VOX_OR / VOX_AND / VOX_ANDNOT   pop b, pop a, push a|b, a&b, a&!b
//   VOX_TRUE / VOX_FALSE            push a constant
//   VOX_TRANSLATE arg = offset of {ox, oy, oz}   save the point, then p - o
//   VOX_SCALE     arg = offset of {sx, sy, sz}   save the point, then p / s
//   VOX_POP                                      restore the saved point
//   (transforms nest at most 8 deep and leave the bit stack alone)
// Program 0 is the air mask; program m >= 1 is material region m, whose id
// (1..255) is written where the air mask and the region both hold.

#include <algorithm>
#include <cstring>
#include <vector>

#include "fdtd_common.hpp"

namespace fdtd_amd {

enum VoxOp : int32_t {
    VOX_FALSE = 0,
    VOX_TRUE = 1,
    VOX_BOX = 2,
    VOX_SPH = 3,
    VOX_CYL = 4,
    VOX_OR = 5,
    VOX_AND = 6,
    VOX_ANDNOT = 7,
    VOX_CONE = 8,
    VOX_TRANSLATE = 9,
    VOX_SCALE = 10,
    VOX_POP = 11,
    VOX_STRICT = 16,
};

constexpr int kVoxMaxDepth = 32;
constexpr int kVoxMaxXform = 8;
constexpr int kVoxMaxProgs = 256;

struct VoxProgram {
    const int32_t* ops;      // device, 2 int32 per instruction, all programs back to back
    const int32_t* start;    // device, n_prog + 1 instruction offsets
    const int32_t* ids;      // device, material id of program m (m >= 1)
    const double* params;    // device
    int n_prog;
};

__device__ __forceinline__ bool cmp_le(double d, bool strict) { return strict ? d < 0.0 : d <= 0.0; }

__device__ __forceinline__ double norm3(double a, double b, double c) {
    return sqrt((a * a + b * b) + c * c);
}

__device__ bool eval_program(const VoxProgram& pg, int m, double x, double y, double z) {
    uint32_t st = 0;  // bit stack, top = bit 0
    double sx[kVoxMaxXform], sy[kVoxMaxXform], sz[kVoxMaxXform];  // saved points
    int xd = 0;
    const int b = pg.start[m], e = pg.start[m + 1];
    for (int q = b; q < e; ++q) {
        const int32_t op = pg.ops[2 * q];
        const int32_t arg = pg.ops[2 * q + 1];
        const bool strict = (op & VOX_STRICT) != 0;
        const int32_t code = op & ~VOX_STRICT;
        if (code == VOX_TRANSLATE || code == VOX_SCALE) {
            const double* P = pg.params + arg;
            sx[xd] = x;
            sy[xd] = y;
            sz[xd] = z;
            ++xd;
            if (code == VOX_TRANSLATE) {
                x = x - P[0];
                y = y - P[1];
                z = z - P[2];
            } else {
                x = x / P[0];
                y = y / P[1];
                z = z / P[2];
            }
            continue;
        }
        if (code == VOX_POP) {
            --xd;
            x = sx[xd];
            y = sy[xd];
            z = sz[xd];
            continue;
        }
        bool bit;
        if (code == VOX_BOX) {
            const double* P = pg.params + arg;
            const double qx = fabs(x - P[0]) - P[3];
            const double qy = fabs(y - P[1]) - P[4];
            const double qz = fabs(z - P[2]) - P[5];
            bit = cmp_le(qx, strict) && cmp_le(qy, strict) && cmp_le(qz, strict);
        } else if (code == VOX_SPH) {
            const double* P = pg.params + arg;
            const double d = norm3(x - P[0], y - P[1], z - P[2]) - P[3];
            bit = cmp_le(d, strict);
        } else if (code == VOX_CYL) {
            const double* P = pg.params + arg;
            const int ax = static_cast<int>(P[8]);
            const double v0 = x - P[0], v1 = y - P[1], v2 = z - P[2];
            const double va = ax == 0 ? v0 : (ax == 1 ? v1 : v2);
            const double na = ax == 0 ? P[3] : (ax == 1 ? P[4] : P[5]);  // +-1
            const double t = na > 0.0 ? va : -va;
            const double len = P[6], r = P[7];
            // np.clip(t, 0, len): min(max(t, 0), len) with numpy's compare-select
            double tc = t > 0.0 ? t : 0.0;
            tc = tc < len ? tc : len;
            // closest = p1 + tc * n: only the axis component moves
            const double c0 = ax == 0 ? P[0] + tc * na : P[0];
            const double c1 = ax == 1 ? P[1] + tc * na : P[1];
            const double c2 = ax == 2 ? P[2] + tc * na : P[2];
            const double radial = norm3(x - c0, y - c1, z - c2);
            const double half = len / 2.0;
            const double ad = fabs(t - half) - half;
            const double axial = 0.0 >= ad ? 0.0 : ad;  // np.maximum(0, ad)
            const bool body = (t >= 0.0) && (t <= len);
            const double rr = radial - r;
            const double d = body ? rr : sqrt(rr * rr + axial * axial);
            bit = cmp_le(d, strict);
        } else if (code == VOX_CONE) {
            const double* P = pg.params + arg;
            const int ax = static_cast<int>(P[9]);
            const double v0 = x - P[0], v1 = y - P[1], v2 = z - P[2];
            const double va = ax == 0 ? v0 : (ax == 1 ? v1 : v2);
            const double na = ax == 0 ? P[3] : (ax == 1 ? P[4] : P[5]);  // +-1
            const double len = P[6], r1 = P[7], r2 = P[8];
            // t = np.dot(v, n) / length; np.clip(t, 0, 1)
            const double t = (na > 0.0 ? va : -va) / len;
            double tc = t > 0.0 ? t : 0.0;
            tc = tc < 1.0 ? tc : 1.0;
            const double rt = r1 + tc * (r2 - r1);
            // closest = p1 + (tc * length) * n: only the axis component moves
            const double s = tc * len;
            const double c0 = ax == 0 ? P[0] + s * na : P[0];
            const double c1 = ax == 1 ? P[1] + s * na : P[1];
            const double c2 = ax == 2 ? P[2] + s * na : P[2];
            const double radial = norm3(x - c0, y - c1, z - c2);
            const double ad = fabs(t - 0.5) - 0.5;
            const double axial = (0.0 >= ad ? 0.0 : ad) * len;  // np.maximum(0, ad) * length
            const bool body = (t >= 0.0) && (t <= 1.0);
            const double rr = radial - rt;
            const double d = body ? rr : sqrt(rr * rr + axial * axial);
            bit = cmp_le(d, strict);
        } else if (code == VOX_TRUE || code == VOX_FALSE) {
            bit = code == VOX_TRUE;
        } else {
            const bool bb = st & 1u, aa = (st >> 1) & 1u;
            st >>= 2;
            bit = code == VOX_OR ? (aa || bb) : (code == VOX_AND ? (aa && bb) : (aa && !bb));
        }
        st = (st << 1) | (bit ? 1u : 0u);
    }
    return st & 1u;
}

__global__ void k_voxelize(const VoxProgram pg, const double* __restrict__ xc, const double* __restrict__ yc,
                           const double* __restrict__ zc, int nx, int ny, int nz, uint8_t* __restrict__ air,
                           uint8_t* __restrict__ ids) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int k = static_cast<int>(c % nz);
        const int j = static_cast<int>((c / nz) % ny);
        const int i = static_cast<int>(c / (static_cast<int64_t>(ny) * nz));
        const double x = xc[i], y = yc[j], z = zc[k];
        const bool a = eval_program(pg, 0, x, y, z);
        uint8_t id = 0;
        if (a)
            for (int m = 1; m < pg.n_prog; ++m)
                if (eval_program(pg, m, x, y, z)) id = static_cast<uint8_t>(pg.ids[m]);
        air[c] = a ? 1 : 0;
        if (ids) ids[c] = id;
    }
}

// Host-side check of a program set: opcodes, parameter ranges, stack depth
// (<= 32, exactly one value left per program), cylinder axes. Returns an
// error message or nullptr.
const char* vox_validate(const int32_t* ops, int n_ops, const int32_t* start, int n_prog, const int32_t* ids,
                         const double* params, int n_params) {
    if (n_prog < 1 || n_prog > kVoxMaxProgs) return "program count must be in [1, 256]";
    if (start[0] != 0 || start[n_prog] != n_ops) return "program offsets must start at 0 and end at n_ops";
    for (int m = 0; m < n_prog; ++m) {
        if (start[m + 1] < start[m]) return "program offsets must not decrease";
        if (m >= 1 && (ids[m] < 1 || ids[m] > 255)) return "material ids must be in [1, 255]";
        int depth = 0, xdepth = 0;
        for (int q = start[m]; q < start[m + 1]; ++q) {
            const int32_t op = ops[2 * q], arg = ops[2 * q + 1];
            const int32_t code = op & ~VOX_STRICT;
            if ((op & ~(VOX_STRICT | 15)) != 0 || code > VOX_POP) return "unknown opcode";
            if (code == VOX_TRANSLATE || code == VOX_SCALE) {
                if (arg < 0 || arg + 3 > n_params) return "parameter offset out of range";
                for (int d = 0; d < 3; ++d) {
                    const double v = params[arg + d];
                    if (code == VOX_SCALE ? !(v >= 1e-6 && v <= 1e6) : !(v == v && v - v == 0.0))
                        return "scale factors must be in [1e-6, 1e6] and offsets finite";
                }
                if (++xdepth > kVoxMaxXform) return "transforms nested deeper than 8";
                continue;
            }
            if (code == VOX_POP) {
                if (--xdepth < 0) return "transform stack underflow";
                continue;
            }
            int need = 0;
            if (code == VOX_BOX) need = 6;
            else if (code == VOX_SPH) need = 4;
            else if (code == VOX_CYL) need = 9;
            else if (code == VOX_CONE) need = 10;
            if (need) {
                if (arg < 0 || arg + need > n_params) return "parameter offset out of range";
                if (code == VOX_CYL || code == VOX_CONE) {
                    const double axd = params[arg + need - 1];
                    if (axd != 0.0 && axd != 1.0 && axd != 2.0) return "cylinder or cone axis must be 0, 1 or 2";
                    const int ax = static_cast<int>(axd);
                    for (int d = 0; d < 3; ++d) {
                        const double nd = params[arg + 3 + d];
                        if (d == ax ? (nd != 1.0 && nd != -1.0) : nd != 0.0)
                            return "cylinder or cone direction must be a signed unit axis";
                    }
                    if (!(params[arg + 6] > 0.0)) return "cylinder or cone length must be positive";
                }
                ++depth;
            } else if (code == VOX_TRUE || code == VOX_FALSE) {
                ++depth;
            } else if (code == VOX_OR || code == VOX_AND || code == VOX_ANDNOT) {
                if (depth < 2) return "stack underflow";
                --depth;
            } else {
                return "unknown opcode";
            }
            if (depth > kVoxMaxDepth) return "stack deeper than 32";
        }
        if (depth != 1) return "a program must leave exactly one value";
        if (xdepth != 0) return "a program must restore every transformed point";
    }
    return nullptr;
}

hipError_t launch_voxelize(const VoxProgram& pg, const double* xc, const double* yc, const double* zc, int nx,
                           int ny, int nz, uint8_t* air, uint8_t* ids, hipStream_t st) {
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    const int blocks = static_cast<int>(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 1 << 16));
    hipLaunchKernelGGL(k_voxelize, dim3(blocks), dim3(256), 0, st, pg, xc, yc, zc, nx, ny, nz, air, ids);
    return hipGetLastError();
}

}  // namespace fdtd_amd
