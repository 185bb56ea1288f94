// Point physics of the DGSEM hot path, device side (gfx950, FP64).
//
// Restates, in the reference's operation order, the two-point fluxes of
// ChandrashekarFlux.cpp:11-161, the log mean of BasicOperations.cpp:6-12, the
// entropy/primitive transforms of BasicOperations.cpp:31-186 and the viscous
// flux of NavierStokesFlux.cpp:12-90.  Normals are axis-aligned on the
// Cartesian meshes this library builds, so a normal is a direction r and a
// positive magnitude; the zero components of the reference's dot products
// are dropped (they add +0 exactly for finite data).
#pragma once
#include <hip/hip_runtime.h>

namespace prandtl {

struct Phys {
    double gamma, gm1, gm1i;
    double mu, lam_h;      // mu and mu*gamma*(1/Pr) (NavierStokesFlux.cpp:38)
    double alpha_max, alpha_min, thresh;
    int sutherland;        // mu(T) per point (NavierStokesFlux.hpp:32-36)
    double R_gas, PrInv;
};

// Viscosity coefficients of one viscous-flux evaluation, in arithmetic R:
// constant mu and lam = mu*gamma*(1/Pr), or Sutherland's law per point.
template <typename R> struct ViscCoef {
    R mu, lam_h;
    bool suth;
    R R_gas, gamma, PrInv;
};

// NavierStokesFlux::ComputeViscosity (NavierStokesFlux.hpp:32-36) with the
// constants of Physics.hpp:35-37; T0pTs = T0 + Ts as in the constructor.
template <typename R>
__device__ __forceinline__ R sutherland_mu(R rho, R p, R R_gas)
{
    const R mu0 = R(1.716e-5), T0 = R(273.15), Ts = R(110.4), T0pTs = R(273.15 + 110.4);
    const R T = p / (rho * R_gas);
    return mu0 * T0pTs / (T + Ts) * (T / T0) * sqrt(T / T0);
}

// Reciprocals: the hardware estimate (v_rcp_f64 / v_rcp_f32) refined by
// Newton-Raphson steps r <- r + r (1 - x r) (two for FP64, one for FP32),
// within an ulp or two of 1/x; the pointwise transforms multiply by them
// where the reference divides.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ float rcp_nr(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// Per-node quantities shared by every two-point flux touching the node.
template <int DIM> struct NodeQ {
    double rho, v[DIM], beta, ibeta, lam;
};

// rho, v = m/rho, p = (gamma-1)(E - rho|v|^2/2), beta = rho/(2p), 1/beta,
// lam = |v| + sqrt(gamma p/rho)  (ChandrashekarFlux.cpp:17-68).  Reciprocals
// of rho and p replace the reference's divisions.
template <int DIM>
__device__ __forceinline__ NodeQ<DIM> node_quants(const double *s, const Phys &P)
{
    NodeQ<DIM> q;
    q.rho = s[0];
    const double ri = rcp_nr(q.rho);
    double vsq = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        q.v[d] = s[1 + d] * ri;
        vsq = (d == 0) ? q.v[d] * q.v[d] : vsq + q.v[d] * q.v[d];
    }
    const double p = P.gm1 * (s[DIM + 1] - 0.5 * q.rho * vsq);
    q.beta = 0.5 * q.rho * rcp_nr(p);
    q.ibeta = 2.0 * p * ri;
    q.lam = sqrt(vsq) + sqrt(P.gamma * p * ri);
    return q;
}

// Logarithmic mean LM(x,y) and its reciprocal, ComputeLogMean of
// BasicOperations.cpp:6-12 (eps = 1e-4).  With d = y-x, s = y+x the
// reference's u = ((xi-1)/(xi+1))^2 with xi = y/x is d^2/s^2, so the branch
// test u < eps is q < eps*t (q = d^2, t = s^2) and the series branch
//   LM = 52.5 s / (105 + 35u + 21u^2 + 15u^3)
//      = 52.5 s t^3 / (((15 q + 21 t) q + 35 t^2) q + 105 t^3)
// costs one division instead of three.  The log branch keeps the reference's
// (y-x)/log(y/x).  Either form differs from the reference's by rounding only
// (t^3 stays in range for |x|,|y| in 1e-40..1e40).
__device__ __forceinline__ double lm_series_den(double q, double t)
{
    return ((15.0 * q + 21.0 * t) * q + 35.0 * t * t) * q + 105.0 * t * t * t;
}

__device__ __forceinline__ double logmean(double x, double y)
{
    const double d = y - x, s = y + x;
    const double q = d * d, t = s * s;
    if (q < 1e-4 * t) return 52.5 * s * (t * t * t) / lm_series_den(q, t);
    return d / log(y / x);
}

__device__ __forceinline__ double inv_logmean(double x, double y)
{
    const double d = y - x, s = y + x;
    const double q = d * d, t = s * s;
    if (q < 1e-4 * t) return lm_series_den(q, t) / (52.5 * s * (t * t * t));
    return log(y / x) / d;
}

// Chandrashekar two-point flux for the normal nr e_r (nr signed) whose
// dissipation is scaled by nd = |nr| (the face normal's Norml2,
// ChandrashekarFlux.cpp:88).  DISS = false: ComputeVolumeFlux
// (ChandrashekarFlux.cpp:11-83), metric mean of two equal metrics.
// DISS = true: ComputeFaceFlux (lines 85-161) with the conserved states s1,
// s2 for the jumps; quirk3d keeps the 3-D energy dissipation term of line
// 131.  Only 1/beta_ln enters the fluxes, so the beta log mean is formed as
// its reciprocal.
template <int DIM, bool DISS>
__device__ __forceinline__ void two_point_flux(int r, double nr, double nd, const NodeQ<DIM> &a,
                                               const NodeQ<DIM> &b, const double *s1,
                                               const double *s2, const Phys &P, bool quirk3d,
                                               double *F)
{
    const double rho_ln = logmean(a.rho, b.rho);
    const double rho_mean = 0.5 * (a.rho + b.rho);
    const double drho = b.rho - a.rho;
    double vm[DIM];
    double h_hat = 0.0, diss = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vm[d] = 0.5 * (a.v[d] + b.v[d]);
        const double hd = -0.25 * (a.v[d] * a.v[d] + b.v[d] * b.v[d]) + vm[d] * vm[d];
        h_hat = (d == 0) ? hd : h_hat + hd;
        if (DISS) {
            const double dv = b.v[d] - a.v[d];
            const double t0 = 0.5 * drho * a.v[d] * b.v[d];
            const double dd = (DIM == 3 && d == 2 && quirk3d) ? t0 + rho_mean * dv + vm[d]
                                                              : t0 + rho_mean * dv * vm[d];
            diss = (d == 0) ? dd : diss + dd;
        }
    }
    double vr = vm[0];
#pragma unroll
    for (int d = 1; d < DIM; d++)
        if (d == r) vr = vm[d];
    const double Vn = vr * nr;
    const double ibeta_ln = inv_logmean(a.beta, b.beta);
    const double p_hat = 0.5 * (a.rho + b.rho) / (a.beta + b.beta);
    h_hat += 0.5 * ibeta_ln * P.gm1i + p_hat / rho_ln;
    const double rVn = rho_ln * Vn;
    F[0] = rVn;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double mom = rho_ln * vm[d];
        F[1 + d] = (d == r) ? Vn * mom + p_hat * nr : Vn * mom;
    }
    F[DIM + 1] = rVn * h_hat;
    if (DISS) {
        diss += 0.5 * drho * P.gm1i * ibeta_ln + 0.5 * rho_mean * P.gm1i * (b.ibeta - a.ibeta);
        const double lam = (a.lam < b.lam) ? b.lam : a.lam;
        const double hl = 0.5 * lam;
        F[0] = F[0] - hl * (s2[0] - s1[0]) * nd;
#pragma unroll
        for (int d = 0; d < DIM; d++) F[1 + d] = F[1 + d] - hl * (s2[1 + d] - s1[1 + d]) * nd;
        F[DIM + 1] = F[DIM + 1] - hl * diss * nd;
    }
}

// MFEM Vector::Norml2 of a 2-vector (scaled, overflow-safe), the face
// normal's magnitude in ChandrashekarFlux::ComputeFaceFlux (line 88).
__device__ __forceinline__ double norml2_2(const double *n)
{
    double scale = 0.0, sum = 0.0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (n[i] != 0.0) {
            const double a = fabs(n[i]);
            if (scale <= a) {
                const double q = scale / a;
                sum = 1.0 + sum * (q * q);
                scale = a;
            } else {
                const double q = a / scale;
                sum += q * q;
            }
        }
    }
    return scale * sqrt(sum);
}

// MFEM Vector::Norml2 of a 3-vector (the same scaled sum).
__device__ __forceinline__ double norml2_3(const double *n)
{
    double scale = 0.0, sum = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (n[i] != 0.0) {
            const double a = fabs(n[i]);
            if (scale <= a) {
                const double q = scale / a;
                sum = 1.0 + sum * (q * q);
                scale = a;
            } else {
                const double q = a / scale;
                sum += q * q;
            }
        }
    }
    return scale * sqrt(sum);
}

template <int DIM> __device__ __forceinline__ double norml2_v(const double *n)
{
    if constexpr (DIM == 2) return norml2_2(n);
    else return norml2_3(n);
}

// two_point_flux for a general metric / normal vector nv (volume pairs:
// the mean of the two nodes' adjugate rows; faces: the face normal), the
// dissipation scaled by nd = |nv|.
template <int DIM, bool DISS>
__device__ __forceinline__ void two_point_flux_v(const double *nv, double nd, const NodeQ<DIM> &a,
                                                 const NodeQ<DIM> &b, const double *s1,
                                                 const double *s2, const Phys &P, bool quirk3d,
                                                 double *F)
{
    const double rho_ln = logmean(a.rho, b.rho);
    const double rho_mean = 0.5 * (a.rho + b.rho);
    const double drho = b.rho - a.rho;
    double vm[DIM];
    double h_hat = 0.0, diss = 0.0, Vn = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vm[d] = 0.5 * (a.v[d] + b.v[d]);
        Vn = (d == 0) ? vm[d] * nv[d] : Vn + vm[d] * nv[d];
        const double hd = -0.25 * (a.v[d] * a.v[d] + b.v[d] * b.v[d]) + vm[d] * vm[d];
        h_hat = (d == 0) ? hd : h_hat + hd;
        if (DISS) {
            const double dv = b.v[d] - a.v[d];
            const double t0 = 0.5 * drho * a.v[d] * b.v[d];
            const double dd = (DIM == 3 && d == 2 && quirk3d) ? t0 + rho_mean * dv + vm[d]
                                                              : t0 + rho_mean * dv * vm[d];
            diss = (d == 0) ? dd : diss + dd;
        }
    }
    const double ibeta_ln = inv_logmean(a.beta, b.beta);
    const double p_hat = 0.5 * (a.rho + b.rho) / (a.beta + b.beta);
    h_hat += 0.5 * ibeta_ln * P.gm1i + p_hat / rho_ln;
    const double rVn = rho_ln * Vn;
    F[0] = rVn;
#pragma unroll
    for (int d = 0; d < DIM; d++) F[1 + d] = Vn * (rho_ln * vm[d]) + p_hat * nv[d];
    F[DIM + 1] = rVn * h_hat;
    if (DISS) {
        diss += 0.5 * drho * P.gm1i * ibeta_ln + 0.5 * rho_mean * P.gm1i * (b.ibeta - a.ibeta);
        const double lam = (a.lam < b.lam) ? b.lam : a.lam;
        const double hl = 0.5 * lam;
        F[0] = F[0] - hl * (s2[0] - s1[0]) * nd;
#pragma unroll
        for (int d = 0; d < DIM; d++) F[1 + d] = F[1 + d] - hl * (s2[1 + d] - s1[1 + d]) * nd;
        F[DIM + 1] = F[DIM + 1] - hl * diss * nd;
    }
}

// Boundary types (PRANDTL_BC_*)
enum { BC_PERIODIC = 0, BC_SLIP = 1, BC_INFLOW = 2, BC_OUTFLOW = 3, BC_SYMMETRY = 4, BC_SPECIFIED = 5,
       BC_NOSLIP = 6 };

// SlipWallBdrFaceIntegrator::ComputeBdrFaceInviscidFlux
// (SlipWallBdrFaceIntegrator.cpp:15-60) for the outward normal ns e_r.  The
// unit normal is ns * (1/|ns|) (Normalize, BasicOperations.hpp:39-42);
// RotateState (BasicOperations.cpp:238-259) of an axis-aligned normal
// reduces to the component permutations below: its tangents Normal(n) and
// n x Normal(n) have unit entries, whose products are exact.  p* is the
// wall pressure of the shock (u_n > 0) or rarefaction (u_n <= 0, base
// clipped at 1e-4) branch; F = (0, p* ns e_r, 0).  Returns the face-frame
// |v| + a.
template <int DIM>
__device__ __forceinline__ double slip_wall_flux(int r, double ns, const double *s, const Phys &P,
                                                 double *F)
{
    const double u = ns * (1.0 / fabs(ns));
    double m[DIM];
    if (DIM == 1) {
        // 1-D: state2(1) * nor(0) (SlipWallBdrFaceIntegrator.cpp:22-30)
        m[0] = s[1] * ns;
    } else if (DIM == 2) {
        if (r == 0) { m[0] = s[1] * u; m[1] = s[2] * u; }
        else { m[0] = s[2] * u; m[1] = -(s[1] * u); }
    } else {
        if (r == 0) { m[0] = s[1] * u; m[1] = s[2]; m[DIM - 1] = s[3] * u; }
        else if (r == 1) { m[0] = s[2] * u; m[1] = -s[1]; m[DIM - 1] = s[3] * u; }
        else { m[0] = s[3] * u; m[1] = -s[1]; m[DIM - 1] = -(s[2] * u); }
    }
    // Conserv2Prim (BasicOperations.cpp:172-186): velocities by division,
    // the pressure's by the reciprocal (ComputePressure, lines 31-37)
    const double rho = s[0];
    const double ri = 1.0 / rho;
    double vel[DIM], vv = 0.0, VV = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = m[d] / rho;
        const double V = m[d] * ri;
        VV = (d == 0) ? V * V : VV + V * V;
        vv = (d == 0) ? vel[d] * vel[d] : vv + vel[d] * vel[d];
    }
    const double p = P.gm1 * (s[DIM + 1] - 0.5 * rho * VV);
    const double an = sqrt(P.gamma * p / rho);
    const double un = vel[0];
    const double gp1 = 1.0 + P.gamma, gp1i = 1.0 / gp1;
    double p_star;
    if (un > 0.0)
        p_star = p + 0.25 * un * gp1 * rho * (un + sqrt(un * un + 8.0 * gp1i / rho * p * (P.gm1 * gp1i + 1.0)));
    else {
        double base = 1.0 + 0.5 * P.gm1 * un / an;
        base = (base < 0.0001) ? 0.0001 : base;
        p_star = p * pow(base, 2.0 * P.gamma * P.gm1i);
    }
#pragma unroll
    for (int v = 0; v < DIM + 2; v++) F[v] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++)
        if (d == r) F[1 + d] = p_star * ns;
    return sqrt(vv) + an;
}

// slip_wall_flux for a general 2-D outward normal n: Normalize, then
// RotateState with Normal(n) = (-n1, n0) (BasicOperations.cpp:211-259).
__device__ __forceinline__ double slip_wall_flux_v2(const double *n, const double *s, const Phys &P, double *F)
{
    const double inv = 1.0 / norml2_2(n);
    const double u0 = n[0] * inv, u1 = n[1] * inv;
    const double m[2] = {s[1] * u0 + s[2] * u1, s[1] * -u1 + s[2] * u0};
    const double rho = s[0];
    const double ri = 1.0 / rho;
    double vel[2], vv = 0.0, VV = 0.0;
#pragma unroll
    for (int d = 0; d < 2; d++) {
        vel[d] = m[d] / rho;
        const double V = m[d] * ri;
        VV = (d == 0) ? V * V : VV + V * V;
        vv = (d == 0) ? vel[d] * vel[d] : vv + vel[d] * vel[d];
    }
    const double p = P.gm1 * (s[3] - 0.5 * rho * VV);
    const double an = sqrt(P.gamma * p / rho);
    const double un = vel[0];
    const double gp1 = 1.0 + P.gamma, gp1i = 1.0 / gp1;
    double p_star;
    if (un > 0.0)
        p_star = p + 0.25 * un * gp1 * rho * (un + sqrt(un * un + 8.0 * gp1i / rho * p * (P.gm1 * gp1i + 1.0)));
    else {
        double base = 1.0 + 0.5 * P.gm1 * un / an;
        base = (base < 0.0001) ? 0.0001 : base;
        p_star = p * pow(base, 2.0 * P.gamma * P.gm1i);
    }
    F[0] = 0.0;
    F[1] = p_star * n[0];
    F[2] = p_star * n[1];
    F[3] = 0.0;
    return sqrt(vv) + an;
}

// boundary_flux for a general 2-D outward normal n (general quadrilateral
// meshes); symmetry mirrors m - 2 (m.u) u with the unit normal u.
__device__ __forceinline__ double boundary_flux_v2(int type, const double *n, const double *s, const double *outer,
                                                   const Phys &P, double *F)
{
    if (type == BC_SLIP || type == BC_NOSLIP) return slip_wall_flux_v2(n, s, P, F);
    const double nd = norml2_2(n);
    double g[4];
#pragma unroll
    for (int v = 0; v < 4; v++) g[v] = (type == BC_INFLOW || type == BC_SPECIFIED) ? outer[v] : s[v];
    if (type == BC_SYMMETRY) {
        const double inv = 1.0 / nd;
        const double u0 = n[0] * inv, u1 = n[1] * inv;
        const double mn = s[1] * u0 + s[2] * u1;
        g[1] = s[1] + -2.0 * mn * u0;
        g[2] = s[2] + -2.0 * mn * u1;
    }
    const NodeQ<2> q1 = node_quants<2>(s, P);
    const NodeQ<2> q2 = node_quants<2>(g, P);
    two_point_flux_v<2, true>(n, nd, q1, q2, s, g, P, false, F);
    return (q1.lam < q2.lam) ? q2.lam : q1.lam;
}

// slip_wall_flux for a general 3-D outward normal n.  RotateState with the
// unit normal u and its two tangents (BasicOperations.cpp:211-259) gives the
// normal momentum m.u; the tangential ones enter only |v|, which the
// rotation keeps, so |m| is used directly.
__device__ __forceinline__ double slip_wall_flux_v3(const double *n, const double *s, const Phys &P, double *F)
{
    const double inv = 1.0 / norml2_3(n);
    const double u0 = n[0] * inv, u1 = n[1] * inv, u2 = n[2] * inv;
    const double mn = s[1] * u0 + s[2] * u1 + s[3] * u2;
    const double rho = s[0];
    const double ri = 1.0 / rho;
    const double mm = s[1] * s[1] + s[2] * s[2] + s[3] * s[3];
    const double vv = mm / (rho * rho);
    const double VV = mm * (ri * ri);
    const double p = P.gm1 * (s[4] - 0.5 * rho * VV);
    const double an = sqrt(P.gamma * p / rho);
    const double un = mn / rho;
    const double gp1 = 1.0 + P.gamma, gp1i = 1.0 / gp1;
    double p_star;
    if (un > 0.0)
        p_star = p + 0.25 * un * gp1 * rho * (un + sqrt(un * un + 8.0 * gp1i / rho * p * (P.gm1 * gp1i + 1.0)));
    else {
        double base = 1.0 + 0.5 * P.gm1 * un / an;
        base = (base < 0.0001) ? 0.0001 : base;
        p_star = p * pow(base, 2.0 * P.gamma * P.gm1i);
    }
    F[0] = 0.0;
    F[1] = p_star * n[0];
    F[2] = p_star * n[1];
    F[3] = p_star * n[2];
    F[4] = 0.0;
    return sqrt(vv) + an;
}

// boundary_flux for a general 3-D outward normal n (hexahedral meshes).
__device__ __forceinline__ double boundary_flux_v3(int type, const double *n, const double *s, const double *outer,
                                                   const Phys &P, bool quirk3d, double *F)
{
    if (type == BC_SLIP || type == BC_NOSLIP) return slip_wall_flux_v3(n, s, P, F);
    const double nd = norml2_3(n);
    double g[5];
#pragma unroll
    for (int v = 0; v < 5; v++) g[v] = (type == BC_INFLOW || type == BC_SPECIFIED) ? outer[v] : s[v];
    if (type == BC_SYMMETRY) {
        const double inv = 1.0 / nd;
        const double u[3] = {n[0] * inv, n[1] * inv, n[2] * inv};
        const double mn = s[1] * u[0] + s[2] * u[1] + s[3] * u[2];
#pragma unroll
        for (int d = 0; d < 3; d++) g[1 + d] = s[1 + d] + -2.0 * mn * u[d];
    }
    const NodeQ<3> q1 = node_quants<3>(s, P);
    const NodeQ<3> q2 = node_quants<3>(g, P);
    two_point_flux_v<3, true>(n, nd, q1, q2, s, g, P, quirk3d, F);
    return (q1.lam < q2.lam) ? q2.lam : q1.lam;
}

template <int DIM>
__device__ __forceinline__ double boundary_flux_vn(int type, const double *n, const double *s, const double *outer,
                                                   const Phys &P, bool quirk3d, double *F)
{
    if constexpr (DIM == 2) return boundary_flux_v2(type, n, s, outer, P, F);
    else return boundary_flux_v3(type, n, s, outer, P, quirk3d, F);
}

// ComputeBdrFaceInviscidFlux of a boundary point (BdrFaceIntegrator.cpp:
// 252-258): the outer state of ComputeOuterInviscidState and the face flux
// with state1 = interior s, normal ns e_r (ns signed outward, nd = |ns|).
// Symmetry (SymmetryBdrFaceIntegrator.cpp) mirrors the normal momentum with
// the unit normal u: m_r - 2 (m_r u) u.  Slip walls replace the flux.
// Returns the point's wave speed.
template <int DIM>
__device__ __forceinline__ double boundary_flux(int type, int r, double ns, double nd, const double *s,
                                                const double *outer, const Phys &P, bool quirk3d,
                                                double *F)
{
    // no-slip walls inherit the slip wall's inviscid flux
    if (type == BC_SLIP || type == BC_NOSLIP) return slip_wall_flux<DIM>(r, ns, s, P, F);
    double g[DIM + 2];
#pragma unroll
    for (int v = 0; v < DIM + 2; v++) g[v] = (type == BC_INFLOW || type == BC_SPECIFIED) ? outer[v] : s[v];
    if (type == BC_SYMMETRY) {
        const double u = ns * (1.0 / nd);
#pragma unroll
        for (int d = 0; d < DIM; d++)
            if (d == r) {
                const double mn = s[1 + d] * u;
                g[1 + d] = s[1 + d] + -2.0 * mn * u;
            }
    }
    const NodeQ<DIM> q1 = node_quants<DIM>(s, P);
    const NodeQ<DIM> q2 = node_quants<DIM>(g, P);
    two_point_flux<DIM, true>(r, ns, nd, q1, q2, s, g, P, quirk3d, F);
    return (q1.lam < q2.lam) ? q2.lam : q1.lam;
}

// Conserv2Entropy, BasicOperations.cpp:81-104 (beta = rho/p)
template <int DIM>
__device__ __forceinline__ void cons2entropy(const double *s, const Phys &P, double *w)
{
    const double r = rcp_nr(s[0]);
    double vel[DIM], vv = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * r;
        vv = (d == 0) ? vel[d] * vel[d] : vv + vel[d] * vel[d];
    }
    const double v_sq = 0.5 * vv;
    const double p = P.gm1 * (s[DIM + 1] - s[0] * v_sq);
    const double beta = s[0] * rcp_nr(p);
    // log(p rho^-gamma) as log p - gamma log rho: two logs instead of a log
    // and a pow (rounding-level difference)
    const double sent = log(p) - P.gamma * log(s[0]);
    w[0] = (P.gamma - sent) * P.gm1i - beta * v_sq;
#pragma unroll
    for (int d = 0; d < DIM; d++) w[1 + d] = beta * vel[d];
    w[DIM + 1] = -beta;
}

// ComputeBdrFaceLiftingFlux of a boundary point for the interior entropy
// variables w.  No-slip wall (NoSlipAdiabWallBdrFaceIntegrator.cpp): the
// wall's (w0, beta V, -beta) minus w, beta = -w_last, V = outer[0..DIM-1].
// Base (BdrFaceIntegrator.cpp:276-282): Entropy2Conserv (BasicOperations.cpp:
// 106-122) of w, the outer state (the state itself for outflow, outer for
// inflow / specified state), Conserv2Entropy, minus w, halved.  quirks:
// Entropy2Conserv's in-place `vel /= beta` leaves velocities in w's momentum
// slots, and those are what is subtracted.
template <int DIM>
__device__ __forceinline__ void lift_flux(int type, const double *w, const double *outer, const Phys &P,
                                          bool quirks, double *f)
{
    constexpr int NV = DIM + 2;
    if (type == BC_NOSLIP) {
        const double v = -w[NV - 1];
        f[0] = w[0] - w[0];
#pragma unroll
        for (int d = 0; d < DIM; d++) f[1 + d] = outer[d] * v - w[1 + d];
        f[NV - 1] = -v - w[NV - 1];
        return;
    }
    double wi[NV], cs[NV], g[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) wi[k] = w[k];
    const double beta = -wi[NV - 1];
    const double ib = 1.0 / beta;
    double vv = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        wi[1 + d] = wi[1 + d] * ib;
        vv = (d == 0) ? wi[1] * wi[1] : vv + wi[1 + d] * wi[1 + d];
    }
    const double sent = P.gamma - (wi[0] + 0.5 * beta * vv) * P.gm1;
    cs[0] = pow(exp(-sent) / beta, P.gm1i);
#pragma unroll
    for (int d = 0; d < DIM; d++) cs[1 + d] = cs[0] * wi[1 + d];
    cs[NV - 1] = cs[0] * (1.0 / beta * P.gm1i + 0.5 * vv);
    double gs[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) gs[k] = (type == BC_OUTFLOW) ? cs[k] : outer[k];
    cons2entropy<DIM>(gs, P, g);
#pragma unroll
    for (int k = 0; k < NV; k++) f[k] = (g[k] - (quirks ? wi[k] : w[k])) * 0.5;
}

// EntropyGrad2PrimGrad, BasicOperations.cpp:126-170, one direction in place,
// in arithmetic R (double, or float in the mixed form).  quirks: the density
// term multiplies velocities (the aliased row, lines 137-138 and 164)
// instead of momenta.
template <int DIM, typename R>
__device__ __forceinline__ void entgrad2primgrad(const R *s, R gm1, R gm1i, bool quirks, R *g)
{
    const R half = 0.5;
    const R r = rcp_nr(s[0]);
    R vel[DIM], vv = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * r;
        vv = (d == 0) ? vel[d] * vel[d] : vv + vel[d] * vel[d];
    }
    const R v_sq = half * vv;
    const R p = gm1 * (s[DIM + 1] - s[0] * v_sq);
    const R KE = v_sq * s[0];
    const R pr = p * r, rp = s[0] * rcp_nr(p);   // p/rho, rho/p
    const R g0 = g[0], gL = g[DIM + 1];
    R mdot = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        g[1 + d] = pr * (g[1 + d] + vel[d] * gL);
        const R md = quirks ? vel[d] : s[1 + d];
        mdot = (d == 0) ? md * g[1 + d] : mdot + md * g[1 + d];
    }
    g[0] = s[0] * g0 - gL * (KE - p * gm1i) + rp * mdot;
    g[DIM + 1] = pr * (g[0] + p * gL);
}

// NavierStokesFlux::ComputeViscousFlux (NavierStokesFlux.cpp:12-90), mu_bulk =
// 2/3, in arithmetic R; constant mu or Sutherland's law (ViscCoef).  dq[c][v] = d(prim_v)/dx_c; out
// fl[v][c], v = 1..nv-1 (the mass row is zero and not stored: fl[v-1][c]).
template <int DIM, typename R>
__device__ __forceinline__ void viscous_flux(const R *s, const R (*dq)[DIM + 2], R gm1, R gm1i,
                                             const ViscCoef<R> &vc, R (*fl)[DIM])
{
    constexpr int NV = DIM + 2;
    const R mu_bulk = R(2) / R(3), two = 2, half = 0.5;
    const R rho = s[0];
    const R rr = rcp_nr(rho);
    R vel[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) vel[d] = s[1 + d] * rr;
    R vv = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) vv = (d == 0) ? vel[d] * vel[d] : vv + vel[d] * vel[d];
    const R p = gm1 * (s[NV - 1] - half * rho * vv);
    R mu = vc.mu, lam_h = vc.lam_h;
    if (vc.suth) {
        // SUTHERLAND build: mu from Conserv2Prim's rho and p, lambda = mu*gamma*(1/Pr)
        mu = sutherland_mu<R>(rho, p, vc.R_gas);
        lam_h = mu * vc.gamma * vc.PrInv;
    }
    R div = dq[0][1];
#pragma unroll
    for (int d = 1; d < DIM; d++) div += dq[d][1 + d];
    const R gr = gm1i * rr, pr = p * rr;
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        const R cvdT = gr * (dq[c][NV - 1] - pr * dq[c][0]);
#pragma unroll
        for (int a = 0; a < DIM; a++) {
            R t;
            if (a == c) t = mu * (two * dq[a][1 + a] - mu_bulk * div);
            else t = mu * (dq[c][1 + a] + dq[a][1 + c]);
            fl[a][c] = t;
        }
        R e = vel[0] * fl[0][c];
#pragma unroll
        for (int a = 1; a < DIM; a++) e = e + vel[a] * fl[a][c];
        fl[NV - 2][c] = e + lam_h * cvdT;
    }
}

} // namespace prandtl
