// Harness that drives the reference's own native kernels, compiled from their
// sources under the reference checkout (see oracle/Makefile). TEST
// INFRASTRUCTURE ONLY: it exists to pin the CPU oracle (fdtd_oracle.cpp) to the
// reference arithmetic, and is never part of the product.
//
// The reference binds these kernels through pybind11 (_kernels/kernels.cpp);
// that module is skipped here (it needs pybind11 and loses ADE registrations,
// see ADEMaterialData in _kernels/include/ade.hpp:44-111). Instead this file
// calls the fdtd:: functions directly and composes them in the order
// core/solver.py:2003-2193 uses (python ordering, which is the only ordering
// that runs ADE).

#include "fdtd_step.hpp"
#include "boundaries.hpp"
#include "pml.hpp"
#include "ade.hpp"
#include "microphones.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// Layout-identical to OrcPole / OrcProblem in fdtd_oracle.cpp.
struct RefPole {
    int32_t material_id, target, lorentz, pad;
    double a, b, d;
    double corr;   // unused: the reference forms its own coefficients
};

struct RefProblem {
    int32_t nx, ny, nz;
    int32_t nonuniform;
    double cv, cp;
    const float* inv_face[3];
    const float* inv_cell[3];
    const uint8_t* air;
    const float* decay[3];
    int32_t n_src;
    const int32_t* src_kind;
    const int64_t* src_a;
    const int64_t* src_b;
    const double* src_wave;
    int32_t wave_len;
    int32_t n_probe;
    const int64_t* probe_idx;
    int32_t n_pole;
    const RefPole* poles;
    const uint8_t* material;
    double ade_inv_dx;
};

// Extra per-material data the reference ADE kernels take from ADEMaterialData.
struct RefMaterial {
    int32_t n;               // entries (index = material id)
    const float* rho_inf;
    const float* K_inf;
    double dt;
};

// ---- single-phase wrappers -------------------------------------------------
void ref_update_velocity(const float* p, float* vx, float* vy, float* vz, int nx, int ny, int nz,
                         float cv) {
    fdtd::update_velocity(p, vx, vy, vz, fdtd::GridShape{nx, ny, nz}, cv);
}

void ref_update_pressure(float* p, const float* vx, const float* vy, const float* vz,
                         const uint8_t* air, int nx, int ny, int nz, float cp) {
    fdtd::update_pressure(p, vx, vy, vz, reinterpret_cast<const bool*>(air),
                          fdtd::GridShape{nx, ny, nz}, cp);
}

static fdtd::NonuniformGridData make_nu(const float* const f[3], const float* const c[3], int nx,
                                        int ny, int nz) {
    fdtd::NonuniformGridData g;
    g.inv_dx_face.assign(f[0], f[0] + nx - 1);
    g.inv_dy_face.assign(f[1], f[1] + ny - 1);
    g.inv_dz_face.assign(f[2], f[2] + nz - 1);
    g.inv_dx_cell.assign(c[0], c[0] + nx);
    g.inv_dy_cell.assign(c[1], c[1] + ny);
    g.inv_dz_cell.assign(c[2], c[2] + nz);
    return g;
}

void ref_update_velocity_nonuniform(const float* p, float* vx, float* vy, float* vz, int nx,
                                    int ny, int nz, float cv_base, const float* ifx,
                                    const float* ify, const float* ifz) {
    const float* f[3] = {ifx, ify, ifz};
    std::vector<float> d0(nx, 1.f), d1(ny, 1.f), d2(nz, 1.f);
    const float* c[3] = {d0.data(), d1.data(), d2.data()};
    auto g = make_nu(f, c, nx, ny, nz);
    fdtd::update_velocity_nonuniform(p, vx, vy, vz, fdtd::GridShape{nx, ny, nz}, g, cv_base);
}

void ref_update_pressure_nonuniform(float* p, const float* vx, const float* vy, const float* vz,
                                    const uint8_t* air, int nx, int ny, int nz, float cp_base,
                                    const float* icx, const float* icy, const float* icz) {
    std::vector<float> d0(nx, 1.f), d1(ny, 1.f), d2(nz, 1.f);
    const float* f[3] = {d0.data(), d1.data(), d2.data()};
    const float* c[3] = {icx, icy, icz};
    auto g = make_nu(f, c, nx, ny, nz);
    fdtd::update_pressure_nonuniform(p, vx, vy, vz, reinterpret_cast<const bool*>(air),
                                     fdtd::GridShape{nx, ny, nz}, g, cp_base);
}

void ref_apply_rigid(float* vx, float* vy, float* vz, const uint8_t* air, int nx, int ny,
                     int nz) {
    auto bc = fdtd::precompute_boundary_cells(reinterpret_cast<const bool*>(air),
                                              fdtd::GridShape{nx, ny, nz});
    fdtd::apply_rigid_boundaries(vx, vy, vz, bc);
}

static fdtd::PMLData make_pml(const float* dx, const float* dy, const float* dz, int nx, int ny,
                              int nz) {
    fdtd::PMLData d;
    if (dx) d.decay_x.assign(dx, dx + nx);
    if (dy) d.decay_y.assign(dy, dy + ny);
    if (dz) d.decay_z.assign(dz, dz + nz);
    return d;
}

void ref_apply_pml_velocity(float* vx, float* vy, float* vz, const float* dx, const float* dy,
                            const float* dz, int nx, int ny, int nz) {
    auto d = make_pml(dx, dy, dz, nx, ny, nz);
    fdtd::apply_pml_velocity(vx, vy, vz, d, fdtd::GridShape{nx, ny, nz});
}

void ref_apply_pml_pressure(float* p, const float* dx, const float* dy, const float* dz, int nx,
                            int ny, int nz) {
    auto d = make_pml(dx, dy, dz, nx, ny, nz);
    fdtd::apply_pml_pressure(p, d, fdtd::GridShape{nx, ny, nz});
}

// The reference's own decay precompute (std::exp in float), for comparison
// with the numpy float32 tables the python PML path actually uses.
void ref_initialize_pml(const float* sx, const float* sy, const float* sz, int nx, int ny, int nz,
                        float dt, float* out_x, float* out_y, float* out_z) {
    auto d = fdtd::initialize_pml(sx, sy, sz, fdtd::GridShape{nx, ny, nz}, dt);
    if (sx) std::memcpy(out_x, d.decay_x.data(), sizeof(float) * nx);
    if (sy) std::memcpy(out_y, d.decay_y.data(), sizeof(float) * ny);
    if (sz) std::memcpy(out_z, d.decay_z.data(), sizeof(float) * nz);
}

void ref_compute_divergence(float* div, const float* vx, const float* vy, const float* vz, int nx,
                            int ny, int nz, float inv_dx) {
    fdtd::compute_divergence(div, vx, vy, vz, fdtd::GridShape{nx, ny, nz}, inv_dx);
}

// ---- ADE phases: the reference's own update / correction kernels ------------
// Pole list layout-identical to fdtd_ade_pole (include/fdtd_mi355x.h).
struct RefAdePole {
    int32_t material_id, target, lorentz, field_index;
    float a, b, d, reserved;
};

static fdtd::ADEMaterialData ade_from_list(const RefAdePole* poles, int n, const float* rho_inf,
                                           const float* K_inf) {
    fdtd::ADEMaterialData d;
    d.rho_inf.assign(256, 0.f);
    d.K_inf.assign(256, 0.f);
    if (rho_inf) d.rho_inf.assign(rho_inf, rho_inf + 256);
    if (K_inf) d.K_inf.assign(K_inf, K_inf + 256);
    for (int q = 0; q < n; ++q) {
        const RefAdePole& p = poles[q];
        if (!p.lorentz) {
            fdtd::ADEMaterialData::DebyePole dp;
            dp.material_id = p.material_id;
            dp.target = p.target;
            dp.coeffs.alpha = p.a;
            dp.coeffs.beta = p.b;
            dp.field_index = p.field_index;
            d.debye_poles.push_back(dp);
            d.n_debye_fields = std::max(d.n_debye_fields, p.field_index + 1);
        } else {
            fdtd::ADEMaterialData::LorentzPole lp;
            lp.material_id = p.material_id;
            lp.target = p.target;
            lp.coeffs.a = p.a;
            lp.coeffs.b = p.b;
            lp.coeffs.d = p.d;
            lp.field_index = p.field_index;
            d.lorentz_poles.push_back(lp);
            d.n_lorentz_fields = std::max(d.n_lorentz_fields, p.field_index + 1);
        }
    }
    return d;
}

void ref_update_ade_density_debye(float* J, const float* p, const uint8_t* mat, const RefAdePole* poles, int n,
                                  int nx, int ny, int nz) {
    fdtd::update_ade_density_debye(J, p, mat, ade_from_list(poles, n, nullptr, nullptr), fdtd::GridShape{nx, ny, nz});
}
void ref_update_ade_modulus_debye(float* J, const float* div, const uint8_t* mat, const RefAdePole* poles, int n,
                                  int nx, int ny, int nz) {
    fdtd::update_ade_modulus_debye(J, div, mat, ade_from_list(poles, n, nullptr, nullptr),
                                   fdtd::GridShape{nx, ny, nz});
}
void ref_update_ade_density_lorentz(float* J, float* Jp, const float* p, const uint8_t* mat, const RefAdePole* poles,
                                    int n, int nx, int ny, int nz) {
    fdtd::update_ade_density_lorentz(J, Jp, p, mat, ade_from_list(poles, n, nullptr, nullptr),
                                     fdtd::GridShape{nx, ny, nz});
}
void ref_update_ade_modulus_lorentz(float* J, float* Jp, const float* div, const uint8_t* mat,
                                    const RefAdePole* poles, int n, int nx, int ny, int nz) {
    fdtd::update_ade_modulus_lorentz(J, Jp, div, mat, ade_from_list(poles, n, nullptr, nullptr),
                                     fdtd::GridShape{nx, ny, nz});
}
void ref_apply_ade_velocity_correction(float* vx, float* vy, float* vz, const float* Jd, const float* Jl,
                                       const uint8_t* mat, const RefAdePole* poles, int n, const float* rho_inf,
                                       float dt, float inv_dx, int nx, int ny, int nz) {
    fdtd::apply_ade_velocity_correction(vx, vy, vz, Jd, Jl, mat, ade_from_list(poles, n, rho_inf, nullptr),
                                        fdtd::GridShape{nx, ny, nz}, dt, inv_dx);
}
void ref_apply_ade_pressure_correction(float* p, const float* Jd, const float* Jl, const uint8_t* mat,
                                       const RefAdePole* poles, int n, const float* K_inf, float dt, int nx, int ny,
                                       int nz) {
    fdtd::apply_ade_pressure_correction(p, Jd, Jl, mat, ade_from_list(poles, n, nullptr, K_inf),
                                        fdtd::GridShape{nx, ny, nz}, dt);
}

// ---- full run in the python step ordering --------------------------------
static fdtd::ADEMaterialData make_ade(const RefProblem& pb, const RefMaterial& mt) {
    fdtd::ADEMaterialData d;
    d.rho_inf.assign(mt.rho_inf, mt.rho_inf + mt.n);
    d.K_inf.assign(mt.K_inf, mt.K_inf + mt.n);
    int nd = 0, nl = 0;
    for (int q = 0; q < pb.n_pole; ++q) {
        const RefPole& p = pb.poles[q];
        if (!p.lorentz) {
            fdtd::ADEMaterialData::DebyePole dp;
            dp.material_id = p.material_id;
            dp.target = p.target;
            dp.coeffs.alpha = static_cast<float>(p.a);
            dp.coeffs.beta = static_cast<float>(p.b);
            dp.field_index = nd++;
            d.debye_poles.push_back(dp);
        } else {
            fdtd::ADEMaterialData::LorentzPole lp;
            lp.material_id = p.material_id;
            lp.target = p.target;
            lp.coeffs.a = static_cast<float>(p.a);
            lp.coeffs.b = static_cast<float>(p.b);
            lp.coeffs.d = static_cast<float>(p.d);
            lp.field_index = nl++;
            d.lorentz_poles.push_back(lp);
        }
    }
    d.n_debye_fields = nd;
    d.n_lorentz_fields = nl;
    return d;
}

int ref_run(const RefProblem* pbp, const RefMaterial* mtp, int n_steps, int step0, float* p,
            float* vx, float* vy, float* vz, float* trace) {
    const RefProblem& pb = *pbp;
    const int nx = pb.nx, ny = pb.ny, nz = pb.nz;
    const fdtd::GridShape sh{nx, ny, nz};
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    const bool* geom = reinterpret_cast<const bool*>(pb.air);
    auto bc = fdtd::precompute_boundary_cells(geom, sh);
    auto pml = make_pml(pb.decay[0], pb.decay[1], pb.decay[2], nx, ny, nz);
    const bool any_pml = pml.has_x() || pml.has_y() || pml.has_z();
    fdtd::NonuniformGridData g;
    if (pb.nonuniform) g = make_nu(pb.inv_face, pb.inv_cell, nx, ny, nz);
    const float cv = static_cast<float>(pb.cv), cp = static_cast<float>(pb.cp);

    fdtd::ADEMaterialData ade;
    std::vector<float> Jd, Jl, Jlp, div;
    if (pb.n_pole) {
        ade = make_ade(pb, *mtp);
        Jd.assign(std::max<int64_t>(1, ade.n_debye_fields * n), 0.f);
        Jl.assign(std::max<int64_t>(1, ade.n_lorentz_fields * n), 0.f);
        Jlp.assign(std::max<int64_t>(1, ade.n_lorentz_fields * n), 0.f);
        div.assign(n, 0.f);
    }
    const float dt = static_cast<float>(mtp ? mtp->dt : 0.0);
    const float inv_dx = static_cast<float>(pb.ade_inv_dx);

    for (int s = 0; s < n_steps; ++s) {
        if (pb.n_pole) {
            if (ade.n_density_debye())
                fdtd::update_ade_density_debye(Jd.data(), p, pb.material, ade, sh);
            if (ade.n_density_lorentz())
                fdtd::update_ade_density_lorentz(Jl.data(), Jlp.data(), p, pb.material, ade, sh);
        }
        if (pb.nonuniform) fdtd::update_velocity_nonuniform(p, vx, vy, vz, sh, g, cv);
        else fdtd::update_velocity(p, vx, vy, vz, sh, cv);
        if (pb.n_pole && (ade.n_density_debye() || ade.n_density_lorentz()))
            fdtd::apply_ade_velocity_correction(vx, vy, vz, Jd.data(), Jl.data(), pb.material, ade,
                                                sh, dt, inv_dx);
        fdtd::apply_rigid_boundaries(vx, vy, vz, bc);
        if (pb.n_pole && (ade.n_modulus_debye() || ade.n_modulus_lorentz())) {
            if (pb.nonuniform) fdtd::compute_divergence_nonuniform(div.data(), vx, vy, vz, sh, g);
            else fdtd::compute_divergence(div.data(), vx, vy, vz, sh, inv_dx);
            if (ade.n_modulus_debye())
                fdtd::update_ade_modulus_debye(Jd.data(), div.data(), pb.material, ade, sh);
            if (ade.n_modulus_lorentz())
                fdtd::update_ade_modulus_lorentz(Jl.data(), Jlp.data(), div.data(), pb.material,
                                                 ade, sh);
        }
        if (pb.nonuniform) fdtd::update_pressure_nonuniform(p, vx, vy, vz, geom, sh, g, cp);
        else fdtd::update_pressure(p, vx, vy, vz, geom, sh, cp);
        if (pb.n_pole && (ade.n_modulus_debye() || ade.n_modulus_lorentz())) {
            fdtd::apply_ade_pressure_correction(p, Jd.data(), Jl.data(), pb.material, ade, sh, dt);
            for (int64_t c = 0; c < n; ++c)
                if (!pb.air[c]) p[c] = 0.f;
        }
        if (any_pml) {
            fdtd::apply_pml_velocity(vx, vy, vz, pml, sh);
            fdtd::apply_pml_pressure(p, pml, sh);
        }
        const int col = step0 + s;
        for (int q = 0; q < pb.n_src; ++q) {
            const double w = pb.src_wave[static_cast<int64_t>(q) * pb.wave_len + col];
            if (pb.src_kind[q] == 0) {
                const int64_t c = pb.src_a[q];
                if (pb.air[c]) p[c] = static_cast<float>(static_cast<double>(p[c]) + w);
            } else {
                const int axis = static_cast<int>(pb.src_a[q]);
                const int idx = static_cast<int>(pb.src_b[q]);
                for (int i = 0; i < nx; ++i)
                    for (int j = 0; j < ny; ++j)
                        for (int k = 0; k < nz; ++k) {
                            const int on = axis == 0 ? i : (axis == 1 ? j : k);
                            if (on != idx) continue;
                            const int64_t c = (static_cast<int64_t>(i) * ny + j) * nz + k;
                            if (pb.air[c]) p[c] = static_cast<float>(static_cast<double>(p[c]) + w);
                        }
            }
        }
        for (int q = 0; q < pb.n_probe; ++q)
            trace[static_cast<int64_t>(s) * pb.n_probe + q] = p[pb.probe_idx[q]];
    }
    return 0;
}

int ref_num_threads(void) {
#if FDTD_HAS_OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
