// fsd_stokes.hpp - C++ mirror of the reference's `Stokes` integration method
// (Stokes.h:43-101) on top of the C ABI in fsd.h.  Header only; link
// libfsd.so.  HOOMD is not involved: the caller owns the device arrays
// (positions float4, director float3, unwrapped positions float4, images int3,
// the 11N Velocity output) and passes them to integrateStepOne.
#ifndef FSD_STOKES_HPP
#define FSD_STOKES_HPP

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>

#include "fsd.h"

namespace fsd {

// device arrays of one particle group (reference layouts, see fsd.h)
struct ParticleArrays {
    float *pos4 = nullptr;        // x, y, z, type
    float *ori3 = nullptr;        // unit director
    float *pos_gb4 = nullptr;     // unwrapped positions
    int32_t *image3 = nullptr;    // periodic images
    float *vel11 = nullptr;       // output: [U, W (6N) | stresslet (5N)]
    const float *force6 = nullptr;     // nullptr: the reference force law
    const float *noise_ang3 = nullptr; // rotational noise, nullptr: none
};

class Stokes {
public:
    // constructor arguments of Stokes::Stokes (Stokes.cc:47-83) plus the box and
    // the time step; defaults as integrate.PSEv3 (integrate.py:40-55)
    Stokes(int N, float Lx, float Ly, float Lz, float T, uint32_t seed = 0, float xi = 0.5f, float error = 1e-3f,
           float dt = 1e-3f, int period = 0, float ndsr = 1e-1f, float kappa = 1.0f / 0.05f, float k_n = 1e4f,
           float beta_AF = 0.f, float epsq = 1e-5f, float sqm_B1 = 0.f, float sqm_B2 = 0.f, uint32_t N_mix = 1,
           float coef_B1_mask = 1.f, float coef_B2_mask = 1.f, float rot_diff = 0.f, float T_ext = 0.f,
           float omega_ext = 0.f)
    {
        p_ = fsd_params{};
        p_.N = N;
        p_.Lx = Lx; p_.Ly = Ly; p_.Lz = Lz;
        p_.T = T; p_.seed = seed; p_.xi = xi; p_.error = error; p_.max_strain = 0.5f; p_.dt = dt;
        p_.period = period;
        p_.ndsr = ndsr; p_.kappa = kappa; p_.k_n = k_n; p_.beta_AF = beta_AF; p_.epsq = epsq;
        p_.sqm_B1 = sqm_B1; p_.sqm_B2 = sqm_B2; p_.N_mix = N_mix;
        p_.coef_B1_mask = coef_B1_mask; p_.coef_B2_mask = coef_B2_mask;
        p_.rot_diff = rot_diff; p_.T_ext = T_ext; p_.omega_ext = omega_ext;
        p_.gmres_tol = error;             // the reference's GMRES tolerance is `error`
        p_.gmres_restart = 50; p_.gmres_maxiter = 1000;
        p_.rlub = 4.0f; p_.rp = 2.1f; p_.r_buff = 0.4f; p_.parity = 1;
    }
    ~Stokes() { if (ctx_) fsd_destroy(ctx_); }
    Stokes(const Stokes &) = delete;
    Stokes &operator=(const Stokes &) = delete;

    // knobs the reference hard-codes; call before AllocateWorkSpaces
    fsd_params &params() { return p_; }

    // ---- set-up, in the reference's order (integrate.py:131-136)
    void setShear(std::function<double(unsigned int)> shear_rate, float max_strain)
    {
        shear_rate_ = std::move(shear_rate);
        p_.max_strain = max_strain;
    }
    void setParams() {}            // derived in fsd_create (Stokes.cc:118-696)
    void setResistanceTable(const float *h_dist1000, const float *h_vals22000)
    {
        dist_ = h_dist1000;
        vals_ = h_vals22000;
    }
    void setFriction(const std::string &type, float h0, const float alpha[6])
    {
        friction_ = type;
        h0_ = h0;
        for (int k = 0; k < 6; ++k) alpha_[k] = alpha ? alpha[k] : 0.f;
    }
    void setSparseMath() {}        // sized in fsd_create (Stokes_SparseMath.cc:15-163)
    void AllocateWorkSpaces()      // Stokes.cc:937-1063: create the context
    {
        check(fsd_create(&p_, &ctx_));
        if (!dist_ || !vals_) throw std::runtime_error("setResistanceTable must precede AllocateWorkSpaces");
        check(fsd_set_resistance_table(ctx_, dist_, vals_));
        check(fsd_set_friction(ctx_, friction_.c_str(), h0_, alpha_));
    }
    void setT(float T)
    {
        p_.T = T;
        if (ctx_) check(fsd_set_temperature(ctx_, T));
    }

    // ---- the step (Stokes::integrateStepOne, Stokes.cc:1197-1683)
    fsd_solve_stats integrateStepOne(unsigned int timestep, ParticleArrays &a)
    {
        if (!ctx_) throw std::runtime_error("AllocateWorkSpaces has not been called");
        const float rate = shear_rate_ ? (float)shear_rate_(timestep) : 0.f;
        fsd_solve_stats st{};
        int rc = fsd_step(ctx_, timestep, rate, a.pos4, a.ori3, a.pos_gb4, a.image3, a.force6, a.noise_ang3,
                          a.vel11, &st);
        check(rc);
        last_converged_ = (rc == FSD_OK);
        return st;
    }
    void integrateStepTwo(unsigned int) {}   // no second half step (Stokes.cc:1688-1690)

    bool converged() const { return last_converged_; }
    fsd_ctx *context() { return ctx_; }

private:
    void check(int rc) const
    {
        if (rc < 0) throw std::runtime_error(std::string("fsd: ") + (ctx_ ? fsd_last_error(ctx_) : "create failed"));
    }
    fsd_params p_{};
    fsd_ctx *ctx_ = nullptr;
    std::function<double(unsigned int)> shear_rate_;
    const float *dist_ = nullptr, *vals_ = nullptr;
    std::string friction_ = "NONE";
    float h0_ = 0.f, alpha_[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool last_converged_ = true;
};

}  // namespace fsd

#endif
