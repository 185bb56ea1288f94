// C++ caller of the drop-in surface (include/prandtl_operator.hpp) on the GPU,
// checked against the CPU oracle (oracle/liboracle.so) on the same input:
// 2-D isentropic vortex, 16x16 quads, N = 3, RHS and three RK3 steps; then
// the autotuned launch shape and a two-case ensemble, each bitwise against
// the default path; the same box as a periodic quadrilateral mesh.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "prandtl_oracle.h"
#include "prandtl_operator.hpp"

static double rel_max(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return num / den;
}

int main()
{
    const int N = 3, Np = N + 1, nx = 16, ny = 16;
    prandtl_mesh_desc mesh = {2, N, {nx, ny, 1}, {-5.0, -5.0, 0.0}, {5.0, 5.0, 1.0}};
    prandtl_params prm = {1.4, 0.72, 0.0, 0.5, 0.001, 0, 0, 0, 1, PRANDTL_FP64};
    hipStream_t stream;
    if (hipStreamCreate(&stream) != hipSuccess) { std::printf("FAIL stream\n"); return 1; }
    prandtl_amd::DGSEMOperator op(mesh, prm, stream);
    prandtl_amd::RK3SSPSolver ode;
    ode.Init(op);

    // isentropic vortex (beta = 5, freestream rho = u = v = p = 1)
    std::vector<double> x(Np), w(Np);
    prandtl_lgl(N, x.data(), w.data());
    const int NE = op.NumElements(), K = op.NodesPerElement();
    const size_t NK = (size_t)NE * K;
    std::vector<double> u(4 * NK);
    const double h = 10.0 / nx, g = 1.4, beta = 5.0, pi = std::acos(-1.0);
    for (int e = 0; e < NE; e++)
        for (int i = 0; i < K; i++) {
            const double X = -5.0 + (e % nx + x[i % Np]) * h, Y = -5.0 + (e / nx + x[i / Np]) * h;
            const double r2 = X * X + Y * Y;
            const double f = beta / (2 * pi) * std::exp(0.5 * (1 - r2));
            const double T = 1 - (g - 1) * beta * beta / (8 * g * pi * pi) * std::exp(1 - r2);
            const double rho = std::pow(T, 1 / (g - 1)), uu = 1 - f * Y, vv = 1 + f * X, p = rho * T;
            u[0 * NK + e * K + i] = rho;
            u[1 * NK + e * K + i] = rho * uu;
            u[2 * NK + e * K + i] = rho * vv;
            u[3 * NK + e * K + i] = p / (g - 1) + 0.5 * rho * (uu * uu + vv * vv);
        }
    double *du = nullptr, *dr = nullptr;
    if (hipMalloc(&du, u.size() * 8) != hipSuccess || hipMalloc(&dr, u.size() * 8) != hipSuccess) {
        std::printf("FAIL alloc\n");
        return 1;
    }
    (void)hipMemcpy(du, u.data(), u.size() * 8, hipMemcpyHostToDevice);
    op.Mult(du, dr);
    std::vector<double> rhs(u.size()), step(u.size());
    (void)hipStreamSynchronize(stream);
    (void)hipMemcpy(rhs.data(), dr, rhs.size() * 8, hipMemcpyDeviceToHost);

    orc_params op_ = {};
    op_.dim = 2; op_.order = N; op_.nelem[0] = nx; op_.nelem[1] = ny; op_.nelem[2] = 1;
    op_.lo[0] = op_.lo[1] = -5.0; op_.hi[0] = op_.hi[1] = 5.0; op_.hi[2] = 1.0;
    op_.gamma = 1.4; op_.Pr = 0.72; op_.alpha_max = 0.5; op_.alpha_min = 0.001; op_.quirks = 1;
    op_.threads = 4;
    orc_ctx *orc = nullptr;
    if (orc_create(&op_, &orc) != 0) { std::printf("FAIL oracle\n"); return 1; }
    std::vector<double> ref(u.size());
    orc_rhs(orc, u.data(), ref.data());
    const double r1 = rel_max(rhs, ref);

    double t = 0.0, dt = 0.01;
    std::vector<double> uref = u;
    for (int s = 0; s < 3; s++) {
        ode.Step(du, t, dt);
        orc_rk3_step(orc, uref.data(), dt);
    }
    (void)hipStreamSynchronize(stream);
    (void)hipMemcpy(step.data(), du, step.size() * 8, hipMemcpyDeviceToHost);
    const double r2 = rel_max(step, uref);
    const double lam = op.GetMaxCharSpeed(), lam_ref = orc_max_wave_speed(orc);
    bool ok = r1 <= 1e-12 && r2 <= 1e-12 && std::fabs(lam - lam_ref) <= 1e-14 * lam_ref &&
              std::fabs(t - 0.03) < 1e-15 && op.StateIsPhysical(du);

    // the autotuned launch shape gives bitwise the same RHS
    (void)hipMemcpy(du, u.data(), u.size() * 8, hipMemcpyHostToDevice);
    const auto shapes = op.Autotune(du);
    op.Mult(du, dr);
    std::vector<double> rhs2(u.size());
    (void)hipStreamSynchronize(stream);
    (void)hipMemcpy(rhs2.data(), dr, rhs2.size() * 8, hipMemcpyDeviceToHost);
    const bool tuned_same = rhs2 == rhs;

    // an ensemble of this case and a coarser N=2 one: bitwise the same as
    // stepping each alone
    prandtl_mesh_desc mesh2 = {2, 2, {8, 6, 1}, {-5.0, -5.0, 0.0}, {5.0, 5.0, 1.0}};
    prandtl_amd::Ensemble ens({mesh, mesh2}, {prm, prm}, stream);
    prandtl_amd::DGSEMOperator op2(mesh2, prm, stream);
    const size_t n2 = 4 * (size_t)op2.NumElements() * op2.NodesPerElement();
    std::vector<double> u2(n2);
    for (size_t i = 0; i < n2; i++) u2[i] = (i < n2 / 4) ? 1.0 + 0.1 * std::sin(0.01 * i) : 0.5 + 0.01 * (i % 7);
    double *ea = nullptr, *eb = nullptr, *sa = nullptr, *sb = nullptr;
    if (hipMalloc(&ea, u.size() * 8) != hipSuccess || hipMalloc(&eb, n2 * 8) != hipSuccess ||
        hipMalloc(&sa, u.size() * 8) != hipSuccess || hipMalloc(&sb, n2 * 8) != hipSuccess) {
        std::printf("FAIL alloc\n");
        return 1;
    }
    (void)hipMemcpy(ea, u.data(), u.size() * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(sa, u.data(), u.size() * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(eb, u2.data(), n2 * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(sb, u2.data(), n2 * 8, hipMemcpyHostToDevice);
    double *ptrs[2] = {ea, eb};
    const double dts[2] = {0.01, 0.02};
    for (int s = 0; s < 2; s++) {
        ens.Step(ptrs, dts);
        op.Step(sa, dts[0]);
        op2.Step(sb, dts[1]);
    }
    (void)hipStreamSynchronize(stream);
    std::vector<double> va(u.size()), vb(u.size()), wa(n2), wb(n2);
    (void)hipMemcpy(va.data(), ea, u.size() * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(vb.data(), sa, u.size() * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(wa.data(), eb, n2 * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(wb.data(), sb, n2 * 8, hipMemcpyDeviceToHost);
    const bool ens_same = va == vb && wa == wb;

    // the same box given as a periodic quadrilateral mesh (per-element
    // corners, wrapped vertices identified): the general-geometry operator
    // gives the box operator's RHS
    std::vector<int> qelem(4 * (size_t)nx * ny);
    std::vector<double> qcoords(8 * (size_t)nx * ny), qvert(2 * (size_t)nx * ny, 0.0);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int e = j * nx + i;
            const int ci[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int k = 0; k < 4; k++) {
                const int a = i + ci[k][0], b = j + ci[k][1];
                qelem[4 * e + k] = (b % ny) * nx + (a % nx);
                qcoords[8 * e + 2 * k] = -5.0 + a * h;
                qcoords[8 * e + 2 * k + 1] = -5.0 + b * h;
            }
        }
    prandtl_quad_mesh qm = {nx * ny, qvert.data(), nx * ny, qelem.data(), qcoords.data(), 0, nullptr, nullptr,
                            0, nullptr, nullptr};
    prandtl_amd::DGSEMOperator opq(qm, N, prm, stream);
    double *dq = nullptr;
    if (hipMalloc(&dq, u.size() * 8) != hipSuccess) { std::printf("FAIL alloc\n"); return 1; }
    opq.Mult(du, dq);
    (void)hipStreamSynchronize(stream);
    std::vector<double> rq(u.size());
    (void)hipMemcpy(rq.data(), dq, rq.size() * 8, hipMemcpyDeviceToHost);
    const double rqrel = rel_max(rq, rhs);
    const bool quad_same = rqrel < 1e-13 && std::fabs(opq.MinElementSize() - op.MinElementSize()) < 1e-15;
    (void)hipFree(dq);

    ok = ok && tuned_same && ens_same && quad_same;
    std::printf("%s rhs_rel=%.3e step_rel=%.3e lambda=%.17g/%.17g shapes=%d/%d tuned_same=%d ensemble_same=%d "
                "quad_rel=%.3e\n",
                ok ? "OK" : "FAIL", r1, r2, lam, lam_ref, shapes.rhs, shapes.visc, (int)tuned_same, (int)ens_same,
                rqrel);
    (void)hipFree(ea);
    (void)hipFree(eb);
    (void)hipFree(sa);
    (void)hipFree(sb);
    orc_destroy(orc);
    (void)hipFree(du);
    (void)hipFree(dr);
    (void)hipStreamDestroy(stream);
    return ok ? 0 : 1;
}
