// prandtl_operator.hpp — header-only C++ surface over the C ABI of
// prandtl_hip.h, shaped like the reference's operator so a C++ caller can
// swap it in for the CPU path on one GPU.
//
//   Prandtl::DGSEMOperator (DGSEMOperator.hpp:15-141)   -> prandtl_amd::DGSEMOperator
//     void Mult(const Vector &u, Vector &dudt) const     -> Mult(const double *u, double *dudt) const
//     real_t GetMaxCharSpeed()                           -> GetMaxCharSpeed()
//   mfem::RK3SSPSolver::Step(x, t, dt)                   -> RK3SSPSolver::Step(double *x, double &t, double &dt)
//
// Pointers are device pointers in the reference's byNODES layout
// (var*NE*K + elem*K + node).  Errors throw prandtl_amd::Error on the C++
// side; nothing throws across the C ABI.  An mfem::TimeDependentOperator
// subclass that forwards Vector::Read()/Write() device pointers to this
// class is the whole of the MFEM-side binding.
#pragma once

#include <stdexcept>
#include <string>

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
        check(prandtl_sizes(ctx_, &NE_, &K_, &nv_), ctx_);
        order_ = mesh.order;
        hmin_ = (mesh.hi[0] - mesh.lo[0]) / mesh.nelem[0];
        for (int r = 1; r < mesh.dim; r++) {
            const double h = (mesh.hi[r] - mesh.lo[r]) / mesh.nelem[r];
            hmin_ = h < hmin_ ? h : hmin_;
        }
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

    // one SSP-RK3 step in place (the three stages fused into the RHS kernels)
    void Step(double *u, double dt) const { check(prandtl_rk3_step(ctx_, u, dt, stream_), ctx_); }

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

    prandtl_ctx *Handle() const { return ctx_; }

  private:
    static void check(int rc, const prandtl_ctx *c)
    {
        if (rc != PRANDTL_OK) throw Error(rc, prandtl_last_error(c));
    }
    prandtl_ctx *ctx_ = nullptr;
    void *stream_ = nullptr;
    int NE_ = 0, K_ = 0, nv_ = 0, order_ = 0;
    double hmin_ = 0.0;
};

// mfem::RK3SSPSolver surface (Simulation.cpp:186-189, 1010)
class RK3SSPSolver {
  public:
    void Init(const DGSEMOperator &f) { f_ = &f; }
    void Step(double *x, double &t, double &dt)
    {
        f_->Step(x, dt);
        t += dt;
    }

  private:
    const DGSEMOperator *f_ = nullptr;
};

} // namespace prandtl_amd
