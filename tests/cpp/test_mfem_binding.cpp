// The MFEM binding of INTEGRATION.md (dgsem_operator_hip.hpp) compiled against
// the MFEM stand-in and driven the way Simulation drives its operator
// (Simulation.cpp:186-189, 807, 1010): DGSEMOperatorHIP::Mult(const Vector&,
// Vector&) and mfem::RK3SSPSolver::Step over it, on a 3-D Navier-Stokes
// Taylor-Green state with a density perturbation (4^3 hexahedra, N = 3,
// Re 1600, Ma 0.1), faithful and corrected, checked against the oracle's RHS
// (1e-12 relative max-norm), its RK3 steps (1e-12) and its running wave
// speed.  Prints "OK" when every check passes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "dgsem_operator_hip.hpp"
#include "prandtl_oracle.h"

static double rel_max(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return num / den;
}

static std::vector<double> to_host(const mfem::Vector &v)
{
    std::vector<double> h(v.Size());
    (void)hipDeviceSynchronize();
    (void)hipMemcpy(h.data(), v.Read(), h.size() * 8, hipMemcpyDeviceToHost);
    return h;
}

static int run(int quirks)
{
    const int N = 3, Np = N + 1, n = 4;
    const double L = 2.0 * std::acos(-1.0), g = 1.4, Ma = 0.1;
    prandtl_mesh_desc mesh = {};
    mesh.dim = 3;
    mesh.order = N;
    for (int r = 0; r < 3; r++) { mesh.nelem[r] = n; mesh.lo[r] = 0.0; mesh.hi[r] = L; }
    prandtl_params prm = {};
    prm.gamma = g; prm.Pr = 0.71; prm.mu = 1.0 / 1600.0;
    prm.alpha_max = 0.5; prm.alpha_min = 0.001;
    prm.parabolic = 1; prm.quirks = quirks; prm.precision = PRANDTL_FP64;

    DGSEMOperatorHIP f(mesh, prm);
    const int size = f.Height();
    const int NE = n * n * n, K = Np * Np * Np;
    const size_t NK = (size_t)NE * K;
    if (size != (int)(5 * NK)) { std::printf("FAIL height %d\n", size); return 1; }

    // Taylor-Green velocity and pressure, density 1 + 0.1 sin x sin y sin z
    std::vector<double> x(Np), w(Np);
    prandtl_lgl(N, x.data(), w.data());
    const double h = L / n, p0 = 1.0 / (g * Ma * Ma);
    std::vector<double> u(size);
    for (int e = 0; e < NE; e++)
        for (int i = 0; i < K; i++) {
            const double X = (e % n + x[i % Np]) * h, Y = ((e / n) % n + x[(i / Np) % Np]) * h,
                         Z = (e / (n * n) + x[i / (Np * Np)]) * h;
            const double rho = 1.0 + 0.1 * std::sin(X) * std::sin(Y) * std::sin(Z);
            const double vx = std::sin(X) * std::cos(Y) * std::cos(Z), vy = -std::cos(X) * std::sin(Y) * std::cos(Z);
            const double p = p0 + (std::cos(2 * X) + std::cos(2 * Y)) * (std::cos(2 * Z) + 2.0) / 16.0;
            const size_t q = (size_t)e * K + i;
            u[q] = rho;
            u[NK + q] = rho * vx;
            u[2 * NK + q] = rho * vy;
            u[3 * NK + q] = 0.0;
            u[4 * NK + q] = p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy);
        }
    mfem::Vector U(size), dUdt(size);
    (void)hipMemcpy(U.Write(), u.data(), u.size() * 8, hipMemcpyHostToDevice);
    f.Mult(U, dUdt);
    const std::vector<double> rhs = to_host(dUdt);

    orc_params op = {};
    op.dim = 3; op.order = N;
    for (int r = 0; r < 3; r++) { op.nelem[r] = n; op.lo[r] = 0.0; op.hi[r] = L; }
    op.gamma = g; op.Pr = 0.71; op.mu = 1.0 / 1600.0; op.parabolic = 1;
    op.alpha_max = 0.5; op.alpha_min = 0.001; op.quirks = quirks; op.threads = 4;
    orc_ctx *orc = nullptr;
    if (orc_create(&op, &orc) != 0) { std::printf("FAIL oracle\n"); return 1; }
    std::vector<double> ref(size);
    orc_rhs(orc, u.data(), ref.data());
    const double r_rhs = rel_max(rhs, ref);
    const double lam_gpu = f.GetMaxCharSpeed(), lam_ref = orc_max_wave_speed(orc);

    // three steps of MFEM's RK3SSPSolver over DGSEMOperatorHIP::Mult
    mfem::RK3SSPSolver ode;
    ode.Init(f);
    double t = 0.0, dt = 2e-3;
    std::vector<double> uref = u;
    for (int s = 0; s < 3; s++) {
        ode.Step(U, t, dt);
        orc_rk3_step(orc, uref.data(), dt);
    }
    const double r_step = rel_max(to_host(U), uref);
    // the solver's last SetTime (its third stage, t + dt/2 of the last step)
    // reached the library
    const double t_lib = f.LibraryTime();
    orc_destroy(orc);
    std::printf("quirks=%d: Mult vs oracle %.3e, 3 RK3SSPSolver steps %.3e, t %.17g, library time %.17g, "
                "wave speed %.17g / %.17g\n", quirks, r_rhs, r_step, t, t_lib, lam_gpu, lam_ref);
    if (!(r_rhs <= 1e-12 && r_step <= 1e-12 && std::fabs(t - 3 * dt) <= 1e-15 &&
          std::fabs(t_lib - (2 * dt + dt / 2)) <= 1e-15 && std::fabs(lam_gpu - lam_ref) <= 1e-14 * lam_ref))
        return 1;
    return 0;
}

int main()
{
    int bad = 0;
    try {
        bad += run(1);
        bad += run(0);
    } catch (const prandtl_amd::Error &e) {
        std::printf("FAIL library error %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
