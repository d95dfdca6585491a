// CPU oracle for the staggered pressure-velocity FDTD step.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP backend in
// strata_fdtd_amd: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it. Nothing in the product path links or calls it.
//
// It restates, in plain C++17 + OpenMP, the arithmetic of the reference native
// kernels and the order in which the reference solver calls them:
//   velocity update      _kernels/src/fdtd_step.cpp:15-82   (nonuniform :235-307)
//   rigid-face zeroing   _kernels/src/boundaries.cpp:13-89
//   pressure update      _kernels/src/fdtd_step.cpp:84-212  (nonuniform :309-439)
//   PML damping          _kernels/src/pml.cpp:47-149, decay tables as
//                        initialize_pml (pml.cpp:13-45) builds them
//   ADE updates          _kernels/src/ade.cpp:25-222, corrections :228-473,
//                        divergence :479-692
//   step ordering        core/solver.py:2003-2077 (python ordering, which is the
//                        only ordering that runs ADE: core/solver.py:2135-2193)
//   sources / probes     core/solver.py:2386-2439
//   boundary planes      first-order Mur ABC and radiation impedance,
//                        boundaries/_boundaries.py:420-527, 529-792
//
// Rounding rules kept on purpose (they decide bit-exact parity):
//   * every expression is evaluated in float with the same association as the
//     reference; the library is built with -ffp-contract=off so no FMA forms;
//   * a point/plane source adds a double waveform value to the float cell and
//     rounds once (numpy float32 + float64 scalar promotes to float64);
//   * PML multiplies p by decay_x, then decay_y, then decay_z (three roundings).
//
// A double-precision twin of the whole step (template parameter T) provides the
// "error vs FP64" figure.
//
// Pinned by (tests/test_oracle.py): bit-exact equality with the reference's own
// kernels compiled from source (oracle/_ref) phase by phase and over whole runs
// on seeded scenes (uniform and nonuniform grids, rigid walls, PML, plane
// sources, Debye and Lorentz density and modulus poles, ADE on a nonuniform
// grid); golden vectors written from those kernels (tests/golden/*.npz); a 1D
// plane wave against a textbook leapfrog; wave speed, cavity mode, p = 0 in
// solids, closed-box energy; Debye fixed point and Lorentz ringing frequency.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline int64_t lin(int i, int j, int k, int ny, int nz) {
    return (static_cast<int64_t>(i) * ny + j) * nz + k;
}

// ---------------------------------------------------------------------------
// Velocity: v[face] += coeff * (p[c+e] - p[c]) on every face below the last
// plane of its axis. The last face of each axis is never touched here.
// Uniform: coeff = cv.  Nonuniform: coeff = cv_base * inv_face[index] formed
// first, exactly as fdtd_step.cpp:262 and :302 do.
// ---------------------------------------------------------------------------
template <typename T>
void velocity(const T* p, T* vx, T* vy, T* vz, int nx, int ny, int nz, T cv,
              const T* ifx, const T* ify, const T* ifz) {
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            if (i < nx - 1) {
                const T c = ifx ? static_cast<T>(cv * ifx[i]) : cv;
                for (int k = 0; k < nz; ++k) vx[b + k] += c * (p[b + sx + k] - p[b + k]);
            }
            if (j < ny - 1) {
                const T c = ify ? static_cast<T>(cv * ify[j]) : cv;
                for (int k = 0; k < nz; ++k) vy[b + k] += c * (p[b + sy + k] - p[b + k]);
            }
            for (int k = 0; k < nz - 1; ++k) {
                const T c = ifz ? static_cast<T>(cv * ifz[k]) : cv;
                vz[b + k] += c * (p[b + k + 1] - p[b + k]);
            }
        }
    }
}

// Rigid faces: a face is zeroed when either of its two cells is solid
// (geometry byte 0). Faces on the last plane of an axis are not part of the
// reference's index lists and are left alone.
template <typename T>
void rigid(T* vx, T* vy, T* vz, const uint8_t* air, int nx, int ny, int nz) {
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            for (int k = 0; k < nz; ++k) {
                const int64_t c = b + k;
                if (i < nx - 1 && (!air[c] || !air[c + sx])) vx[c] = T(0);
                if (j < ny - 1 && (!air[c] || !air[c + sy])) vy[c] = T(0);
                if (k < nz - 1 && (!air[c] || !air[c + 1])) vz[c] = T(0);
            }
        }
    }
}

// Pressure: solid cells are set to 0; air cells get
//   uniform:    p += cp * ((dvx + dvy) + dvz)
//   nonuniform: p += cp_base * ((icx*dvx + icy*dvy) + icz*dvz)
// with the face below index 0 of each axis read as 0.
template <typename T>
void pressure(T* p, const T* vx, const T* vy, const T* vz, const uint8_t* air,
              int nx, int ny, int nz, T cp, const T* icx, const T* icy, const T* icz) {
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            for (int k = 0; k < nz; ++k) {
                const int64_t c = b + k;
                if (!air[c]) { p[c] = T(0); continue; }
                const T dvx = (i > 0) ? T(vx[c] - vx[c - sx]) : vx[c];
                const T dvy = (j > 0) ? T(vy[c] - vy[c - sy]) : vy[c];
                const T dvz = (k > 0) ? T(vz[c] - vz[c - 1]) : vz[c];
                if (icx) {
                    const T ddx = icx[i] * dvx, ddy = icy[j] * dvy, ddz = icz[k] * dvz;
                    p[c] += cp * T(T(ddx + ddy) + ddz);
                } else {
                    p[c] += cp * T(T(dvx + dvy) + dvz);
                }
            }
        }
    }
}

// PML (native pml.cpp form): each table may be null (axis without PML).
template <typename T>
void pml_velocity(T* vx, T* vy, T* vz, const T* dx, const T* dy, const T* dz,
                  int nx, int ny, int nz) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            for (int k = 0; k < nz; ++k) {
                if (dx) vx[b + k] *= dx[i];
                if (dy) vy[b + k] *= dy[j];
                if (dz) vz[b + k] *= dz[k];
            }
        }
    }
}

template <typename T>
void pml_pressure(T* p, const T* dx, const T* dy, const T* dz, int nx, int ny, int nz) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            for (int k = 0; k < nz; ++k) {
                T v = p[b + k];
                if (dx) v *= dx[i];
                if (dy) v *= dy[j];
                if (dz) v *= dz[k];
                p[b + k] = v;
            }
        }
    }
}

// Divergence used as the modulus-pole source (native ade.cpp:479-692):
//   div = dvx*ix; div += dvy*iy; div += dvz*iz   (in that order)
// Uniform grids pass ix = iy = iz = 1/dx through the *_u scalars.
template <typename T>
void divergence(T* div, const T* vx, const T* vy, const T* vz, int nx, int ny, int nz,
                T inv_u, const T* icx, const T* icy, const T* icz) {
    const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int64_t b = lin(i, j, 0, ny, nz);
            const T ix = icx ? icx[i] : inv_u;
            const T iy = icy ? icy[j] : inv_u;
            for (int k = 0; k < nz; ++k) {
                const int64_t c = b + k;
                const T iz = icz ? icz[k] : inv_u;
                const T dvx = (i > 0) ? T(vx[c] - vx[c - sx]) : vx[c];
                const T dvy = (j > 0) ? T(vy[c] - vy[c - sy]) : vy[c];
                const T dvz = (k > 0) ? T(vz[c] - vz[c - 1]) : vz[c];
                T d = dvx * ix;
                d += dvy * iy;
                d += dvz * iz;
                div[c] = d;
            }
        }
    }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Problem description shared with the Python side of the oracle (ctypes).
// ---------------------------------------------------------------------------
struct OrcPole {
    int32_t material_id;   // 1..255
    int32_t target;        // 0 density, 1 modulus
    int32_t lorentz;       // 0 Debye, 1 Lorentz
    int32_t pad;
    double a, b, d;        // Debye: alpha=a, beta=b.  Lorentz: a, b, d
    double corr;           // density: -dt/rho_inf * (1/min spacing)
                           // modulus: -K_inf*dt
};

struct OrcProblem {
    int32_t nx, ny, nz;
    int32_t nonuniform;            // 0: scalar cv/cp; 1: per-axis tables below
    double cv, cp;                 // uniform coeff_v/coeff_p, or the *_base ones
    const float* inv_face[3];      // nonuniform: n-1 entries each (else null)
    const float* inv_cell[3];      // nonuniform: n entries each (else null)
    const uint8_t* air;            // 1 = air, 0 = solid (n cells)
    const float* decay[3];         // PML decay per axis or null
    // sources
    int32_t n_src;
    const int32_t* src_kind;       // 0 point, 1 plane
    const int64_t* src_a;          // point: flat index; plane: axis
    const int64_t* src_b;          // plane: index along axis
    const double* src_wave;        // [n_src][wave_len] waveform at t_n
    int32_t wave_len;
    // probes
    int32_t n_probe;
    const int64_t* probe_idx;      // flat indices
    // ADE
    int32_t n_pole;
    const OrcPole* poles;
    const uint8_t* material;       // per-cell material id (null if no poles)
    double ade_inv_dx;             // uniform grids: 1/dx for the divergence source
    // membrane sources (src_kind 2, src_a = field 0 p / 1 vx / 2 vy / 3 vz):
    // per source a float64 weight per cell, null for point and plane sources
    const double* const* src_weight;
    // boundary planes, applied in order after the PML and before the sources
    int32_t n_bnd;
    const struct OrcBoundary* bnd;
    void* bnd_state;               // previous-step planes (float or double, as the run),
                                   // kept by the caller across calls
};

// One boundary plane operation on p:
//   kind 0 (Mur ABC, _boundaries.py:479-511), both faces of `axis`:
//     p[0] = prev0 + coeff (p[1] - p[0]);      prev0 = p[1]
//     p[N-1] = prev1 + coeff (p[N-2] - p[N-1]); prev1 = p[N-2]
//   kind 1 (radiation impedance, :707-776), face `side` (0 low, 1 high):
//     abc = prev + coeff (p_in - p_b); p_b = refl p_in + omr abc; prev = p_in
// Types as numpy forms them: the Mur coefficient is a float64 numpy scalar
// (the solver's dt comes from np.sqrt), so a float32 difference times it is
// float64 and the sums with it are float64, rounded once when stored; R is a
// Python float, rounded to the field's type in refl * p_in; omr = 1 - R in
// double multiplies the float64 abc.
struct OrcBoundary {
    int32_t kind, axis, side, pad;
    double coeff, refl, omr;
    int64_t state_off;             // first element of this entry's planes in bnd_state
};

// Per-phase entry points (float) so tests can compare single phases.
void orc_velocity_f32(const float* p, float* vx, float* vy, float* vz, int nx, int ny, int nz,
                      float cv, const float* ifx, const float* ify, const float* ifz) {
    velocity<float>(p, vx, vy, vz, nx, ny, nz, cv, ifx, ify, ifz);
}
void orc_rigid_f32(float* vx, float* vy, float* vz, const uint8_t* air, int nx, int ny, int nz) {
    rigid<float>(vx, vy, vz, air, nx, ny, nz);
}
void orc_pressure_f32(float* p, const float* vx, const float* vy, const float* vz,
                      const uint8_t* air, int nx, int ny, int nz, float cp,
                      const float* icx, const float* icy, const float* icz) {
    pressure<float>(p, vx, vy, vz, air, nx, ny, nz, cp, icx, icy, icz);
}
void orc_pml_velocity_f32(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                          const float* dz, int nx, int ny, int nz) {
    pml_velocity<float>(vx, vy, vz, dx, dy, dz, nx, ny, nz);
}
void orc_pml_pressure_f32(float* p, const float* dx, const float* dy, const float* dz,
                          int nx, int ny, int nz) {
    pml_pressure<float>(p, dx, dy, dz, nx, ny, nz);
}
void orc_divergence_f32(float* div, const float* vx, const float* vy, const float* vz,
                        int nx, int ny, int nz, float inv_u,
                        const float* icx, const float* icy, const float* icz) {
    divergence<float>(div, vx, vy, vz, nx, ny, nz, inv_u, icx, icy, icz);
}

// ---- ADE phases (float), restating ade.cpp:25-473 -----------------------------
// Pole list: {material_id, target, lorentz, field_index, a, b, d, reserved};
// J arrays hold one full grid per field_index. Each call visits the poles of
// its list (Debye / Lorentz) and target in list order, one pass per pole.
struct OrcAdePole {
    int32_t material_id, target, lorentz, field_index;
    float a, b, d, reserved;
};

static void orc_ade_update(float* J, float* Jp, const float* src, const uint8_t* m, const OrcAdePole* poles, int n,
                           int lorentz, int target, int64_t cells) {
    for (int q = 0; q < n; ++q) {
        const OrcAdePole& pl = poles[q];
        if (pl.lorentz != lorentz || pl.target != target) continue;
        float* Jq = J + pl.field_index * cells;
        float* Jpq = lorentz ? Jp + pl.field_index * cells : nullptr;
        const uint8_t mid = static_cast<uint8_t>(pl.material_id);
        for (int64_t c = 0; c < cells; ++c) {
            if (m[c] != mid) continue;
            if (!lorentz) {
                Jq[c] = pl.a * Jq[c] + pl.b * src[c];
            } else {
                const float j0 = Jq[c], j1 = Jpq[c];
                const float jn = pl.a * j0 + pl.b * j1 + pl.d * src[c];
                Jpq[c] = j0;
                Jq[c] = jn;
            }
        }
    }
}

void orc_update_ade_density_debye(float* J, const float* p, const uint8_t* m, const OrcAdePole* poles, int n, int nx,
                                  int ny, int nz) {
    orc_ade_update(J, nullptr, p, m, poles, n, 0, 0, static_cast<int64_t>(nx) * ny * nz);
}
void orc_update_ade_modulus_debye(float* J, const float* div, const uint8_t* m, const OrcAdePole* poles, int n, int nx,
                                  int ny, int nz) {
    orc_ade_update(J, nullptr, div, m, poles, n, 0, 1, static_cast<int64_t>(nx) * ny * nz);
}
void orc_update_ade_density_lorentz(float* J, float* Jp, const float* p, const uint8_t* m, const OrcAdePole* poles,
                                    int n, int nx, int ny, int nz) {
    orc_ade_update(J, Jp, p, m, poles, n, 1, 0, static_cast<int64_t>(nx) * ny * nz);
}
void orc_update_ade_modulus_lorentz(float* J, float* Jp, const float* div, const uint8_t* m, const OrcAdePole* poles,
                                    int n, int nx, int ny, int nz) {
    orc_ade_update(J, Jp, div, m, poles, n, 1, 1, static_cast<int64_t>(nx) * ny * nz);
}

// Density poles, Debye list then Lorentz list: coeff = -dt / rho_inf * inv_dx
// (float), v[face] += coeff * (J[c+e] - J[c]) where both cells hold the material.
void orc_apply_ade_velocity_correction(float* vx, float* vy, float* vz, const float* Jd, const float* Jl,
                                       const uint8_t* m, const OrcAdePole* poles, int n, const float* rho_inf,
                                       float dt, float inv_dx, int nx, int ny, int nz) {
    const int64_t cells = static_cast<int64_t>(nx) * ny * nz, sy = nz, sx = static_cast<int64_t>(ny) * nz;
    for (int lorentz = 0; lorentz < 2; ++lorentz)
        for (int q = 0; q < n; ++q) {
            const OrcAdePole& pl = poles[q];
            if (pl.lorentz != lorentz || pl.target != 0) continue;
            const float* J = (lorentz ? Jl : Jd) + pl.field_index * cells;
            const uint8_t mid = static_cast<uint8_t>(pl.material_id);
            const float coeff = -dt / rho_inf[mid] * inv_dx;
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < nz; ++k) {
                        const int64_t c = lin(i, j, k, ny, nz);
                        if (m[c] != mid) continue;
                        if (i < nx - 1 && m[c + sx] == mid) vx[c] += coeff * (J[c + sx] - J[c]);
                        if (j < ny - 1 && m[c + sy] == mid) vy[c] += coeff * (J[c + sy] - J[c]);
                        if (k < nz - 1 && m[c + 1] == mid) vz[c] += coeff * (J[c + 1] - J[c]);
                    }
        }
}

// Modulus poles, Debye list then Lorentz list: p += (-K_inf * dt) * J.
void orc_apply_ade_pressure_correction(float* p, const float* Jd, const float* Jl, const uint8_t* m,
                                       const OrcAdePole* poles, int n, const float* K_inf, float dt, int nx, int ny,
                                       int nz) {
    const int64_t cells = static_cast<int64_t>(nx) * ny * nz;
    for (int lorentz = 0; lorentz < 2; ++lorentz)
        for (int q = 0; q < n; ++q) {
            const OrcAdePole& pl = poles[q];
            if (pl.lorentz != lorentz || pl.target != 1) continue;
            const float* J = (lorentz ? Jl : Jd) + pl.field_index * cells;
            const uint8_t mid = static_cast<uint8_t>(pl.material_id);
            const float coeff = -K_inf[mid] * dt;
            for (int64_t c = 0; c < cells; ++c)
                if (m[c] == mid) p[c] += coeff * J[c];
        }
}

// PML decay as the native initialize_pml forms it: std::exp(-sigma * dt) with
// float sigma and float dt (_kernels/src/pml.cpp:13-45).
void orc_pml_decay(const float* sigma, int n, float dt, float* out) {
    for (int i = 0; i < n; ++i) out[i] = std::exp(-sigma[i] * dt);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

namespace {

// One reference time step in the python ordering (core/solver.py:2135-2193,
// then boundaries :2044-2047, sources :2386-2433, probes :2435-2439).
// J arrays: per pole, a full-grid J (and J_prev for Lorentz).
template <typename T>
struct Runner {
    const OrcProblem& pb;
    int nx, ny, nz;
    int64_t n;
    std::vector<T> ifx, ify, ifz, icx, icy, icz, ddx, ddy, ddz, div;
    std::vector<std::vector<T>> J, Jp;
    T cv, cp, inv_u;

    explicit Runner(const OrcProblem& p) : pb(p), nx(p.nx), ny(p.ny), nz(p.nz) {
        n = static_cast<int64_t>(nx) * ny * nz;
        auto cp_tab = [](const float* src, int m, std::vector<T>& dst) {
            if (!src) return;
            dst.assign(src, src + m);
        };
        if (pb.nonuniform) {
            cp_tab(pb.inv_face[0], nx - 1, ifx); cp_tab(pb.inv_face[1], ny - 1, ify);
            cp_tab(pb.inv_face[2], nz - 1, ifz);
            cp_tab(pb.inv_cell[0], nx, icx); cp_tab(pb.inv_cell[1], ny, icy);
            cp_tab(pb.inv_cell[2], nz, icz);
        }
        cp_tab(pb.decay[0], nx, ddx); cp_tab(pb.decay[1], ny, ddy); cp_tab(pb.decay[2], nz, ddz);
        cv = static_cast<T>(pb.cv);
        cp = static_cast<T>(pb.cp);
        inv_u = static_cast<T>(pb.ade_inv_dx);
        J.resize(pb.n_pole); Jp.resize(pb.n_pole);
        for (int q = 0; q < pb.n_pole; ++q) {
            J[q].assign(n, T(0));
            if (pb.poles[q].lorentz) Jp[q].assign(n, T(0));
        }
        if (pb.n_pole) div.assign(n, T(0));
    }
    const T* tab(const std::vector<T>& v) const { return v.empty() ? nullptr : v.data(); }

    // J <- a J + b src (Debye) or J <- a J + b Jprev + d src (Lorentz) on the
    // cells of the pole's material, ade.cpp:25-222.
    void pole_update(int q, const T* src) {
        const OrcPole& pl = pb.poles[q];
        T* Jq = J[q].data();
        const T a = static_cast<T>(pl.a), b = static_cast<T>(pl.b), d = static_cast<T>(pl.d);
        const uint8_t mid = static_cast<uint8_t>(pl.material_id);
        const uint8_t* m = pb.material;
        if (!pl.lorentz) {
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < n; ++c)
                if (m[c] == mid) Jq[c] = a * Jq[c] + b * src[c];
        } else {
            T* Jpq = Jp[q].data();
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < n; ++c)
                if (m[c] == mid) {
                    const T j0 = Jq[c], j1 = Jpq[c];
                    const T jn = a * j0 + b * j1 + d * src[c];
                    Jpq[c] = j0;
                    Jq[c] = jn;
                }
        }
    }

    // v[face] += corr * (J[c+e] - J[c]) where both cells hold the pole's
    // material, ade.cpp:228-401. corr = -dt/rho_inf * inv_dx with the scalar
    // inv_dx = 1/min spacing the solver passes on every grid, nonuniform ones
    // included (core/solver.py:3295-3305): no per-face spacing here.
    void velocity_correction(int q, T* vx, T* vy, T* vz) {
        const OrcPole& pl = pb.poles[q];
        const T* Jq = J[q].data();
        const uint8_t mid = static_cast<uint8_t>(pl.material_id);
        const uint8_t* m = pb.material;
        const T corr = static_cast<T>(pl.corr);
        const int64_t sy = nz, sx = static_cast<int64_t>(ny) * nz;
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const int64_t b = lin(i, j, 0, ny, nz);
                for (int k = 0; k < nz; ++k) {
                    const int64_t c = b + k;
                    if (m[c] != mid) continue;
                    if (i < nx - 1 && m[c + sx] == mid) vx[c] += corr * (Jq[c + sx] - Jq[c]);
                    if (j < ny - 1 && m[c + sy] == mid) vy[c] += corr * (Jq[c + sy] - Jq[c]);
                    if (k < nz - 1 && m[c + 1] == mid) vz[c] += corr * (Jq[c + 1] - Jq[c]);
                }
            }
    }

    void pressure_correction(int q, T* p) {
        const OrcPole& pl = pb.poles[q];
        const T* Jq = J[q].data();
        const uint8_t mid = static_cast<uint8_t>(pl.material_id);
        const uint8_t* m = pb.material;
        const T corr = static_cast<T>(pl.corr);
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n; ++c)
            if (m[c] == mid) p[c] += corr * Jq[c];
    }

    // Native pole iteration order: Debye poles first, then Lorentz poles, each
    // in registration order (ade.cpp loops debye_poles then lorentz_poles).
    template <typename F>
    void for_poles(int target, F f) {
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < pb.n_pole; ++q)
                if (pb.poles[q].target == target && pb.poles[q].lorentz == pass) f(q);
    }

    // Boundary plane b on p, one cross-section cell at a time: the cells of a
    // cross-section column along the axis are independent of the others, so
    // the reference's whole-plane numpy statements are restated per column in
    // their order.
    void boundary_plane(const OrcBoundary& bd, T* p) {
        const int dims[3] = {nx, ny, nz};
        const int ax = bd.axis, N = dims[ax];
        const int nu = ax == 0 ? ny : nx, nw = ax == 2 ? ny : nz;
        T* st = static_cast<T*>(pb.bnd_state) + bd.state_off;
        const double coeff = bd.coeff, omr = bd.omr;
        const T refl = static_cast<T>(bd.refl);
        auto cell = [&](int a, int u, int w) -> int64_t {
            if (ax == 0) return lin(a, u, w, ny, nz);
            if (ax == 1) return lin(u, a, w, ny, nz);
            return lin(u, w, a, ny, nz);
        };
#pragma omp parallel for collapse(2) schedule(static)
        for (int u = 0; u < nu; ++u)
            for (int w = 0; w < nw; ++w) {
                const int64_t t = static_cast<int64_t>(u) * nw + w;
                if (bd.kind == 0) {
                    T* prev0 = st;
                    T* prev1 = st + static_cast<int64_t>(nu) * nw;
                    const int64_t c0 = cell(0, u, w), c1 = cell(1, u, w);
                    const int64_t cl = cell(N - 1, u, w), cm = cell(N - 2, u, w);
                    const T d0 = p[c1] - p[c0];
                    p[c0] = static_cast<T>(static_cast<double>(prev0[t]) + coeff * static_cast<double>(d0));
                    prev0[t] = p[c1];
                    const T d1 = p[cm] - p[cl];
                    p[cl] = static_cast<T>(static_cast<double>(prev1[t]) + coeff * static_cast<double>(d1));
                    prev1[t] = p[cm];
                } else {
                    const int64_t cb = cell(bd.side ? N - 1 : 0, u, w);
                    const int64_t ci = cell(bd.side ? N - 2 : 1, u, w);
                    const T pin = p[ci];
                    const T d = pin - p[cb];
                    const double abc = static_cast<double>(st[t]) + coeff * static_cast<double>(d);
                    const T rig = refl * pin;
                    p[cb] = static_cast<T>(static_cast<double>(rig) + omr * abc);
                    st[t] = pin;
                }
            }
    }

    void step(T* p, T* vx, T* vy, T* vz, int s, float* trace_row) {
        const bool ade = pb.n_pole > 0;
        if (ade) for_poles(0, [&](int q) { pole_update(q, p); });
        velocity<T>(p, vx, vy, vz, nx, ny, nz, cv, tab(ifx), tab(ify), tab(ifz));
        if (ade) for_poles(0, [&](int q) { velocity_correction(q, vx, vy, vz); });
        rigid<T>(vx, vy, vz, pb.air, nx, ny, nz);
        if (ade) {
            bool any_mod = false;
            for (int q = 0; q < pb.n_pole; ++q) any_mod |= (pb.poles[q].target == 1);
            if (any_mod) {
                divergence<T>(div.data(), vx, vy, vz, nx, ny, nz, inv_u, tab(icx), tab(icy), tab(icz));
                for_poles(1, [&](int q) { pole_update(q, div.data()); });
            }
        }
        // Pressure: the python ordering adds the modulus correction before it
        // zeroes solids; the native kernel zeroes solids first. Both end with
        // p == 0 in solids, and on air cells the two additions are the same.
        pressure<T>(p, vx, vy, vz, pb.air, nx, ny, nz, cp, tab(icx), tab(icy), tab(icz));
        if (ade) {
            for_poles(1, [&](int q) { pressure_correction(q, p); });
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < n; ++c)
                if (!pb.air[c]) p[c] = T(0);
        }
        pml_velocity<T>(vx, vy, vz, tab(ddx), tab(ddy), tab(ddz), nx, ny, nz);
        pml_pressure<T>(p, tab(ddx), tab(ddy), tab(ddz), nx, ny, nz);
        // boundary planes in list order (core/solver.py:2044-2047)
        for (int b = 0; b < pb.n_bnd; ++b) boundary_plane(pb.bnd[b], p);
        // sources in list order (core/solver.py:2386-2433)
        for (int q = 0; q < pb.n_src; ++q) {
            const double w = pb.src_wave[static_cast<int64_t>(q) * pb.wave_len + s];
            if (pb.src_kind[q] == 2) {
                // membrane: field[mask] += w * weight[mask], mask = (weight > 0) & air;
                // numpy adds the float64 product to the element in float64
                T* fld[4] = {p, vx, vy, vz};
                T* F = fld[pb.src_a[q]];
                const double* wt = pb.src_weight[q];
#pragma omp parallel for schedule(static)
                for (int64_t c = 0; c < n; ++c)
                    if (wt[c] > 0.0 && pb.air[c]) F[c] = static_cast<T>(static_cast<double>(F[c]) + w * wt[c]);
                continue;
            }
            if (pb.src_kind[q] == 0) {
                const int64_t c = pb.src_a[q];
                if (pb.air[c]) p[c] = static_cast<T>(static_cast<double>(p[c]) + w);
            } else {
                const int axis = static_cast<int>(pb.src_a[q]);
                const int idx = static_cast<int>(pb.src_b[q]);
#pragma omp parallel for collapse(2) schedule(static)
                for (int a0 = 0; a0 < (axis == 0 ? ny : nx); ++a0)
                    for (int a1 = 0; a1 < (axis == 2 ? ny : nz); ++a1) {
                        int i, j, k;
                        if (axis == 0) { i = idx; j = a0; k = a1; }
                        else if (axis == 1) { i = a0; j = idx; k = a1; }
                        else { i = a0; j = a1; k = idx; }
                        const int64_t c = lin(i, j, k, ny, nz);
                        if (pb.air[c]) p[c] = static_cast<T>(static_cast<double>(p[c]) + w);
                    }
            }
        }
        for (int q = 0; q < pb.n_probe; ++q)
            trace_row[q] = static_cast<float>(p[pb.probe_idx[q]]);
    }
};

}  // namespace

extern "C" {

// Run n_steps from the given state (in place). trace: [n_steps][n_probe].
// step0 selects the waveform column of the first step. Returns 0.
// The problem's boundary planes alone, once, on p (float32, in place), with
// the problem's persistent plane state.
int orc_boundary_planes_f32(const OrcProblem* pb, float* p) {
    Runner<float> r(*pb);
    for (int b = 0; b < pb->n_bnd; ++b) r.boundary_plane(pb->bnd[b], p);
    return 0;
}

int orc_run_f32(const OrcProblem* pb, int n_steps, int step0, float* p, float* vx, float* vy,
                float* vz, float* trace) {
    Runner<float> r(*pb);
    for (int s = 0; s < n_steps; ++s)
        r.step(p, vx, vy, vz, step0 + s, trace + static_cast<int64_t>(s) * pb->n_probe);
    return 0;
}

// Same step sequence in double precision (tables are widened from float).
int orc_run_f64(const OrcProblem* pb, int n_steps, int step0, double* p, double* vx, double* vy,
                double* vz, float* trace) {
    Runner<double> r(*pb);
    for (int s = 0; s < n_steps; ++s)
        r.step(p, vx, vy, vz, step0 + s, trace + static_cast<int64_t>(s) * pb->n_probe);
    return 0;
}

// Time only: run n_steps and return; used by the CPU baseline.
int orc_time_steps_f32(const OrcProblem* pb, int n_steps, float* p, float* vx, float* vy,
                       float* vz, float* trace) {
    return orc_run_f32(pb, n_steps, 0, p, vx, vy, vz, trace);
}

}  // extern "C"
