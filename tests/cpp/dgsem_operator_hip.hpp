// The MFEM-side binding of the library, as INTEGRATION.md gives it: an
// mfem::TimeDependentOperator whose Mult forwards MFEM's device pointers to
// the library (DGSEMOperator::Mult, DGSEMOperator.hpp:105).  Compiled against
// the MFEM stand-in of tests/cpp/mfem_standin by test_mfem_binding.cpp.
#pragma once
#include "mfem.hpp"
#include "prandtl_operator.hpp"

class DGSEMOperatorHIP : public mfem::TimeDependentOperator {
    prandtl_amd::DGSEMOperator op;
public:
    DGSEMOperatorHIP(const prandtl_mesh_desc &m, const prandtl_params &p)
        : mfem::TimeDependentOperator(0), op(m, p) {
        height = width = op.Height();
    }
    void Mult(const mfem::Vector &u, mfem::Vector &dudt) const override {
        op.Mult(u.Read(true), dudt.Write(true));
    }
    // RK3SSPSolver sets each stage's time before its Mult
    void SetTime(const mfem::real_t t_) override {
        mfem::TimeDependentOperator::SetTime(t_);
        op.SetTime(t_);
    }
    double GetMaxCharSpeed() { return op.GetMaxCharSpeed(); }
    double LibraryTime() const { return op.GetTime(); }
};
