// prandtl_operator.hpp — header-only C++ surface over the C ABI of
// prandtl_hip.h, shaped like the reference's operator so a C++ caller can
// swap it in for the CPU path on one GPU.
//
//   Prandtl::DGSEMOperator (DGSEMOperator.hpp:15-141)   -> prandtl_amd::DGSEMOperator
//     void Mult(const Vector &u, Vector &dudt) const     -> Mult(const double *u, double *dudt) const
//     real_t GetMaxCharSpeed()                           -> GetMaxCharSpeed()
//   TimeDependentOperator::SetTime / GetTime           -> SetTime(double) / GetTime()
//   mfem::RK3SSPSolver::Step(x, t, dt)                   -> RK3SSPSolver::Step(double *x, double &t, double &dt)
//   Simulation::Run time loop (Simulation.cpp:1007-1022)  -> DGSEMOperator::Run (on the device)
//   (no counterpart)                                     -> DGSEMOperator::Autotune / SetShape, Ensemble
//
// Pointers are device pointers in the reference's byNODES layout
// (var*NE*K + elem*K + node).  Errors throw prandtl_amd::Error on the C++
// side; nothing throws across the C ABI.  An mfem::TimeDependentOperator
// subclass that forwards Vector::Read()/Write() device pointers to this
// class is the whole of the MFEM-side binding.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "prandtl_hip.h"

namespace prandtl_amd {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

class DGSEMOperator {
  public:
    DGSEMOperator(const prandtl_mesh_desc &mesh, const prandtl_params &params, void *stream = nullptr)
        : stream_(stream)
    {
        check(prandtl_create(&mesh, &params, &ctx_), nullptr);
        init(mesh.order);
    }
    // on a quadrilateral mesh (Simulation.cpp:271-346's mesh set-up; axisymmetric
    // with params.axisymmetric)
    DGSEMOperator(const prandtl_quad_mesh &mesh, int order, const prandtl_params &params, void *stream = nullptr)
        : stream_(stream)
    {
        check(prandtl_create_quad(&mesh, order, &params, &ctx_), nullptr);
        init(order);
    }
    // on a hexahedral mesh
    DGSEMOperator(const prandtl_hex_mesh &mesh, int order, const prandtl_params &params, void *stream = nullptr)
        : stream_(stream)
    {
        check(prandtl_create_hex(&mesh, order, &params, &ctx_), nullptr);
        init(order);
    }
    ~DGSEMOperator() { prandtl_destroy(ctx_); }
    DGSEMOperator(const DGSEMOperator &) = delete;
    DGSEMOperator &operator=(const DGSEMOperator &) = delete;

    // mfem::Operator::Height()/Width(): nv * NE * K
    long Height() const { return (long)nv_ * NE_ * K_; }
    long Width() const { return Height(); }
    int NumElements() const { return NE_; }
    int NodesPerElement() const { return K_; }
    int NumEquations() const { return nv_; }
    int Order() const { return order_; }
    double MinElementSize() const { return hmin_; }

    // DGSEMOperator::Mult: dudt is overwritten with L(u)
    void Mult(const double *u, double *dudt) const { check(prandtl_rhs(ctx_, u, dudt, stream_), ctx_); }

    // one SSP-RK3 step in place (the three stages fused into the RHS kernels),
    // its stages at t, t + dt and t + dt/2 of the operator's time, which it
    // then advances by dt
    void Step(double *u, double dt) const { check(prandtl_rk3_step(ctx_, u, dt, stream_), ctx_); }

    // TimeDependentOperator::SetTime / GetTime (DGSEMOperator.hpp:111-114):
    // the time at which Mult evaluates time-dependent boundary states
    void SetTime(double t) const { check(prandtl_set_time(ctx_, t, stream_), ctx_); }
    double GetTime() const
    {
        double t = 0.0;
        check(prandtl_get_time(ctx_, &t), ctx_);
        return t;
    }

    // boundary attribute attr's outer state: below where
    // normal . x - offset - speed t < 0, above elsewhere (2-D Euler quadrilateral meshes)
    void SetBoundarySplit(int attr, const double normal[2], double offset, double speed, const double *below,
                          const double *above)
    {
        check(prandtl_set_boundary_split(ctx_, attr, normal, offset, speed, below, above), ctx_);
    }

    // algorithm of the RHS kernel (PRANDTL_ALGO_TILE or PRANDTL_ALGO_LINE)
    void SetAlgorithm(int algo) { check(prandtl_set_algorithm(ctx_, PRANDTL_KERNEL_RHS, algo), ctx_); }

    // Simulation::Run's time loop on the device (prandtl_run): u from t to
    // t_final with variable dt from cfl (> 0) or the fixed dt; returns steps
    int Run(double *u, double &t, double t_final, double cfl, double dt = 0.0, int max_steps = 0,
            int check_every = 0) const
    {
        int steps = 0;
        check(prandtl_run(ctx_, u, &t, t_final, cfl, dt, max_steps, check_every, &steps, stream_), ctx_);
        return steps;
    }

    // DGSEMOperator::GetMaxCharSpeed: running maximum of |v| + c
    double GetMaxCharSpeed(bool reset = false) const
    {
        double v = 0.0;
        check(prandtl_max_wave_speed(ctx_, &v, reset ? 1 : 0, stream_), ctx_);
        return v;
    }

    bool StateIsPhysical(const double *u) const
    {
        const int rc = prandtl_check_state(ctx_, u, stream_);
        if (rc == PRANDTL_ERR_STATE) return false;
        check(rc, ctx_);
        return true;
    }

    // launch shapes of the two kernels, timed on state u (prandtl_autotune)
    struct Shapes {
        int rhs = 0, visc = 0;
    };
    Shapes Autotune(const double *u)
    {
        int sh[2] = {0, 0};
        check(prandtl_autotune(ctx_, u, sh, nullptr, stream_), ctx_);
        return Shapes{sh[0], sh[1]};
    }
    void SetShape(int kernel, int shape) { check(prandtl_set_shape(ctx_, kernel, shape), ctx_); }

    // RecoverStateFromWeighted (axisymmetric operators): U of the weighted state rU
    void RecoverState(const double *rU, double *U) const
    {
        check(prandtl_axis_recover(ctx_, rU, U, stream_), ctx_);
    }

    prandtl_ctx *Handle() const { return ctx_; }

  private:
    static void check(int rc, const prandtl_ctx *c)
    {
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_last_error(c));
    }
    void init(int order)
    {
        check(prandtl_sizes(ctx_, &NE_, &K_, &nv_), ctx_);
        check(prandtl_hmin(ctx_, &hmin_), ctx_);
        order_ = order;
    }
    prandtl_ctx *ctx_ = nullptr;
    void *stream_ = nullptr;
    int NE_ = 0, K_ = 0, nv_ = 0, order_ = 0;
    double hmin_ = 0.0;
};

// Independent 2-D Euler cases stepped together, one launch per RK stage
// (prandtl_ensemble_*).  Case contexts belong to the ensemble.
class Ensemble {
  public:
    Ensemble(const std::vector<prandtl_mesh_desc> &meshes, const std::vector<prandtl_params> &params,
             void *stream = nullptr)
        : stream_(stream)
    {
        if (meshes.size() != params.size()) throw Error(PRANDTL_ERR_ARG, "meshes and params differ in length");
        const int rc = prandtl_ensemble_create((int)meshes.size(), meshes.data(), params.data(), &ens_);
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_ensemble_last_error(nullptr));
    }
    ~Ensemble() { prandtl_ensemble_destroy(ens_); }
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;

    // one SSP-RK3 step of case i's state u[i] by dt[i], for every case
    void Step(double *const *u, const double *dt)
    {
        const int rc = prandtl_ensemble_rk3_step(ens_, u, dt, stream_);
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_ensemble_last_error(ens_));
    }
    // launch form of the stages: 0 one workgroup per tile, g > 0 persistent
    // with g workgroups per CU; Autotune times them on u and keeps the fastest
    void SetMode(int mode)
    {
        const int rc = prandtl_ensemble_set_mode(ens_, mode);
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_ensemble_last_error(ens_));
    }
    int Autotune(double *const *u, const double *dt)
    {
        int mode = 0;
        const int rc = prandtl_ensemble_autotune(ens_, u, dt, &mode, nullptr, stream_);
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_ensemble_last_error(ens_));
        return mode;
    }
    prandtl_ctx *Case(int i)
    {
        prandtl_ctx *c = nullptr;
        const int rc = prandtl_ensemble_case(ens_, i, &c);
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_ensemble_last_error(ens_));
        return c;
    }

  private:
    prandtl_ensemble *ens_ = nullptr;
    void *stream_ = nullptr;
};

// mfem::RK3SSPSolver surface (Simulation.cpp:186-189, 1010)
class RK3SSPSolver {
  public:
    void Init(const DGSEMOperator &f) { f_ = &f; }
    void Step(double *x, double &t, double &dt)
    {
        f_->SetTime(t);
        f_->Step(x, dt);
        t += dt;
    }

  private:
    const DGSEMOperator *f_ = nullptr;
};

} // namespace prandtl_amd
