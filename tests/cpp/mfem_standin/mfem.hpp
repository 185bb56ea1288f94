// A minimal stand-in for the part of MFEM's API that the binding of
// INTEGRATION.md uses, with MFEM's names and signatures (linalg/vector.hpp:
// Vector with Read/Write/ReadWrite device access, Add, add; linalg/operator.hpp:
// Operator, TimeDependentOperator; linalg/ode.hpp: ODESolver, RK3SSPSolver),
// so that the binding compiles and runs here, where MFEM is not installed.
// Test infrastructure: a Vector here lives in device memory only (MFEM built
// with MFEM_USE_HIP keeps a host copy too), and the vector updates run as
// small HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace mfem {

using real_t = double;

namespace standin {
inline void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "mfem stand-in: %s: %s\n", what, hipGetErrorString(e));
        std::abort();
    }
}
__global__ void k_axpby(int n, real_t a, const real_t *x, real_t b, const real_t *y, real_t *z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}
__global__ void k_addto(int n, real_t a, const real_t *x, real_t *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
inline int blocks(int n) { return (n + 255) / 256; }
} // namespace standin

class Vector {
  public:
    Vector() = default;
    explicit Vector(int n) { SetSize(n); }
    Vector(const Vector &v) : Vector(v.size_) { *this = v; }
    ~Vector() { if (data_) (void)hipFree(data_); }
    void SetSize(int n)
    {
        if (data_) (void)hipFree(data_);
        data_ = nullptr;
        size_ = n;
        if (n > 0) standin::hip_ok(hipMalloc(&data_, sizeof(real_t) * n), "hipMalloc");
    }
    int Size() const { return size_; }
    // device pointers (on_dev is MFEM's "use the device" flag; only true is meaningful here)
    const real_t *Read(bool on_dev = true) const { (void)on_dev; return data_; }
    real_t *Write(bool on_dev = true) { (void)on_dev; return data_; }
    real_t *ReadWrite(bool on_dev = true) { (void)on_dev; return data_; }
    Vector &operator=(const Vector &v)
    {
        if (size_ != v.size_) SetSize(v.size_);
        standin::hip_ok(hipMemcpy(data_, v.data_, sizeof(real_t) * size_, hipMemcpyDeviceToDevice), "copy");
        return *this;
    }
    // this += a * Va
    Vector &Add(const real_t a, const Vector &Va)
    {
        hipLaunchKernelGGL(standin::k_addto, dim3(standin::blocks(size_)), dim3(256), 0, 0, size_, a, Va.data_,
                           data_);
        return *this;
    }

  private:
    int size_ = 0;
    real_t *data_ = nullptr;
};

// z = a x + b y and z = x + a y (MFEM's free functions add())
inline void add(const real_t a, const Vector &x, const real_t b, const Vector &y, Vector &z)
{
    hipLaunchKernelGGL(standin::k_axpby, dim3(standin::blocks(x.Size())), dim3(256), 0, 0, x.Size(), a, x.Read(),
                       b, y.Read(), z.Write());
}
inline void add(const Vector &x, const real_t a, const Vector &y, Vector &z) { add(1.0, x, a, y, z); }

class Operator {
  public:
    explicit Operator(int s = 0) : height(s), width(s) {}
    virtual ~Operator() = default;
    int Height() const { return height; }
    int Width() const { return width; }
    virtual void Mult(const Vector &x, Vector &y) const = 0;

  protected:
    int height, width;
};

class TimeDependentOperator : public Operator {
  public:
    explicit TimeDependentOperator(int n = 0, real_t t_ = 0.0) : Operator(n), t(t_) {}
    virtual real_t GetTime() const { return t; }
    virtual void SetTime(const real_t t_) { t = t_; }

  protected:
    real_t t;
};

class ODESolver {
  public:
    virtual ~ODESolver() = default;
    virtual void Init(TimeDependentOperator &f_) { f = &f_; }
    virtual void Step(Vector &x, real_t &t, real_t &dt) = 0;

  protected:
    TimeDependentOperator *f = nullptr;
};

// Third-order SSP Runge-Kutta in the Shu-Osher form, one Operator::Mult per
// stage: y = x + dt f(x); y = 3/4 x + 1/4 (y + dt f(y)); x = 1/3 x + 2/3 (y +
// dt f(y)).
class RK3SSPSolver : public ODESolver {
  public:
    void Init(TimeDependentOperator &f_) override
    {
        ODESolver::Init(f_);
        y.SetSize(f->Width());
        k.SetSize(f->Width());
    }
    void Step(Vector &x, real_t &t, real_t &dt) override
    {
        f->SetTime(t);
        f->Mult(x, k);
        add(x, dt, k, y);
        f->SetTime(t + dt);
        f->Mult(y, k);
        y.Add(dt, k);
        add(3.0 / 4.0, x, 1.0 / 4.0, y, y);
        f->SetTime(t + dt / 2);
        f->Mult(y, k);
        y.Add(dt, k);
        add(1.0 / 3.0, x, 2.0 / 3.0, y, x);
        t += dt;
    }

  private:
    Vector y, k;
};

} // namespace mfem
