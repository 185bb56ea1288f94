/*
 * TEST INFRASTRUCTURE — not part of the product.
 *
 * CPU FP64 restatement ("the oracle") of Prandtl's DGSEM right-hand side,
 * its BR1 lifting, modal indicator with subcell-FV blending, and MFEM's
 * SSP-RK3 step.  It is the checker that the HIP path in prandtl_amd/ is
 * compared against; the product never links or calls it.
 *
 * Every routine cites the reference lines it restates (paths relative to the
 * reference tree).  The arithmetic keeps the reference's operation order and
 * association wherever the reference fixes it (for example MFEM's
 * `Vector /= c` multiplies by 1/c, `a*b*c` is (a*b)*c), so that the oracle and
 * the reference differ only where MFEM internals that are not in the tree
 * (LGL node rounding, dense inverse rounding) fix the last bits.
 *
 * All of it is templated on the scalar type: `double` for values and `Cnt`,
 * an op-counting scalar, for the algorithmic flop count (each +,-,*,/, sqrt,
 * log, exp, pow counts one).
 *
 * Geometry is general per-node (J, adjugate rows, face normals, subcell
 * normals) so the restatement follows the reference's metric use; it is
 * built for Cartesian boxes (periodic or with boundary sides; 1-D segments
 * with the Euler equations included), quadrilateral meshes (bilinear) and
 * hexahedral meshes (trilinear), and the axisymmetric mode recovers U from
 * r U (RecoverStateFromWeighted).
 *
 * Pins (tests/): each function is checked against closed forms, invariants
 * or equivalences that a plausible slip breaks — test_oracle.py (fluxes,
 * operators, RK3), test_oracle_bc.py and test_oracle_ns_walls.py (boundary
 * fluxes and lifting), test_oracle_quad.py and test_oracle_hex.py (mesh
 * geometry: equality with the box operator, free stream, conservation,
 * rotation and relabelling), test_oracle_axisym.py (recovery closed forms,
 * exact steady states, conservation, convergence), test_oracle_1d.py (1-D:
 * equality with the y-uniform 2-D box, free stream, conservation, the 1-D
 * slip wall, exact Riemann solutions, convergence).  Parity unpinned: the
 * last bits that MFEM's own LGL rounding, dense inverse and RK3 axpy order
 * would fix (MFEM is not in the reference tree).
 */
#include "prandtl_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------ */
/* op-counting scalar                                                   */
/* ------------------------------------------------------------------ */
long long g_ops[8];
int g_cat = 0;

struct Cnt {
    double v;
    Cnt() : v(0.0) {}
    Cnt(double x) : v(x) {}
};
/* Books the operations that follow to category k.  Only the counting
 * instantiation, which runs on one thread, sets it: the other scalar types
 * leave the shared variable alone, so their threads do not write it. */
template <class T> inline void set_cat(int k)
{
    if constexpr (std::is_same<T, Cnt>::value) g_cat = k;
    else (void)k;
}
inline Cnt operator+(Cnt a, Cnt b) { ++g_ops[g_cat]; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_ops[g_cat]; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_ops[g_cat]; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_ops[g_cat]; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt &operator+=(Cnt &a, Cnt b) { a = a + b; return a; }
inline Cnt &operator-=(Cnt &a, Cnt b) { a = a - b; return a; }
inline Cnt &operator*=(Cnt &a, Cnt b) { a = a * b; return a; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator!=(Cnt a, Cnt b) { return a.v != b.v; }

inline double val(double x) { return x; }
inline double val(Cnt x) { return x.v; }
inline double val(long double x) { return (double)x; }
inline long double mlog(long double x) { return std::log(x); }
inline long double msqrt(long double x) { return std::sqrt(x); }
inline long double mexp(long double x) { return std::exp(x); }
inline long double mpow(long double x, double y) { return std::pow(x, (long double)y); }
inline long double mabs(long double x) { return std::fabs(x); }
inline double mlog(double x) { return std::log(x); }
inline Cnt mlog(Cnt x) { ++g_ops[g_cat]; return Cnt(std::log(x.v)); }
inline double msqrt(double x) { return std::sqrt(x); }
inline Cnt msqrt(Cnt x) { ++g_ops[g_cat]; return Cnt(std::sqrt(x.v)); }
inline double mexp(double x) { return std::exp(x); }
inline Cnt mexp(Cnt x) { ++g_ops[g_cat]; return Cnt(std::exp(x.v)); }
inline double mpow(double x, double y) { return std::pow(x, y); }
inline Cnt mpow(Cnt x, double y) { ++g_ops[g_cat]; return Cnt(std::pow(x.v, y)); }
inline double mabs(double x) { return std::fabs(x); }
inline Cnt mabs(Cnt x) { return Cnt(std::fabs(x.v)); }
/* std::max semantics: max(a,b) = (a < b) ? b : a */
template <class T> inline T mmax(T a, T b) { return (a < b) ? b : a; }

enum { CAT_TOTAL = 0, CAT_IND = 1, CAT_LIFT = 2, CAT_VISC = 3, CAT_VOL = 4,
       CAT_FACE = 5, CAT_FV = 6, CAT_ASM = 7 };

/* ------------------------------------------------------------------ */
/* 1-D rules (DGSEMIntegrator.cpp:18-82, ModalBasis.cpp:110-135)         */
/* ------------------------------------------------------------------ */

/* Legendre-Gauss-Lobatto nodes and weights on [0,1]: the rule MFEM returns
 * for IntegrationRules(0, GaussLobatto).Get(SEGMENT, 2Np-3)
 * (DGSEMIntegrator.cpp:18-22).  Interior nodes are the roots of P_N',
 * found by Newton's method in long double from Chebyshev-Lobatto guesses;
 * weights 2/(N(N+1) P_N(x)^2) on [-1,1], halved on [0,1]. */
void lgl_rule(int N, std::vector<double> &x, std::vector<double> &w)
{
    const int Np = N + 1;
    x.assign(Np, 0.0);
    w.assign(Np, 0.0);
    if (N == 0) { x[0] = 0.5; w[0] = 1.0; return; }
    std::vector<long double> z(Np), wz(Np);
    for (int i = 0; i < Np; i++) {
        long double t = -std::cos((long double)M_PI * i / N);
        if (i == 0) t = -1.0L;
        else if (i == N) t = 1.0L;
        else {
            for (int it = 0; it < 100; it++) {
                /* P_N, P_N' by recurrence */
                long double p0 = 1.0L, p1 = t, dp0 = 0.0L, dp1 = 1.0L;
                for (int n = 1; n < N; n++) {
                    long double p2 = ((2 * n + 1) * t * p1 - n * p0) / (n + 1);
                    long double dp2 = dp0 + (2 * n + 1) * p1;
                    p0 = p1; p1 = p2; dp0 = dp1; dp1 = dp2;
                }
                long double PN = (N == 1) ? t : p1;
                long double dPN = (N == 1) ? 1.0L : dp1;
                long double d2PN = (2.0L * t * dPN - N * (N + 1) * PN) / (1.0L - t * t);
                long double dt = dPN / d2PN;
                t -= dt;
                if (std::fabs((double)dt) < 1e-22) break;
            }
        }
        long double p0 = 1.0L, p1 = t;
        for (int n = 1; n < N; n++) {
            long double p2 = ((2 * n + 1) * t * p1 - n * p0) / (n + 1);
            p0 = p1; p1 = p2;
        }
        long double PN = (N == 1) ? t : p1;
        z[i] = t;
        wz[i] = 2.0L / (N * (N + 1) * PN * PN);
    }
    /* symmetric pairs share magnitudes; the middle node of odd Np is 1/2 */
    for (int i = 0; i < Np; i++) {
        long double zs = 0.5L * (z[i] - z[N - i]);
        long double ws = 0.5L * (wz[i] + wz[N - i]);
        x[i] = (double)(0.5L * (1.0L + zs));
        w[i] = (double)(0.5L * ws);
    }
    x[0] = 0.0;
    x[N] = 1.0;
    if (Np % 2 == 1) x[N / 2] = 0.5;
}

/* D, Dhat = D + E, Dhat2 = 2D + E with E = diag(+1/w_0, 0, ..., -1/w_N);
 * barycentric weights and negative-sum diagonal exactly as
 * DGSEMIntegrator.cpp:45-80 (stored here un-transposed, row-major). */
void dmats(int N, const std::vector<double> &x, const std::vector<double> &w,
           std::vector<double> &D, std::vector<double> &Dh, std::vector<double> &Dh2)
{
    const int Np = N + 1;
    std::vector<double> wb(Np, 1.0);
    for (int i = 1; i < Np; i++)
        for (int j = 0; j < i; j++) {
            wb[j] *= (x[j] - x[i]);
            wb[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i < Np; i++) wb[i] = 1.0 / wb[i];
    D.assign(Np * Np, 0.0);
    for (int iL = 0; iL < Np; iL++)
        for (int i = 0; i < Np; i++)
            if (iL != i) {
                D[i * Np + iL] = wb[iL] / wb[i] / (x[i] - x[iL]);
                D[i * Np + i] -= D[i * Np + iL];
            }
    Dh = D;
    Dh[0] += 1.0 / w[0];
    Dh[N * Np + N] -= 1.0 / w[N];
    Dh2.assign(Np * Np, 0.0);
    for (int i = 0; i < Np * Np; i++) Dh2[i] = D[i] * 2.0;
    Dh2[0] += 1.0 / w[0];
    Dh2[N * Np + N] -= 1.0 / w[N];
}

/* shifted Legendre P_n(2x-1), MFEM Poly_1D::CalcLegendre (ModalBasis.cpp:123) */
void calc_legendre(int p, double x, double *u)
{
    u[0] = 1.0;
    if (p == 0) return;
    double z = 2.0 * x - 1.0;
    u[1] = z;
    for (int n = 1; n < p; n++) u[n + 1] = ((2 * n + 1) * z * u[n] - n * u[n - 1]) / (n + 1);
}

/* dense inverse by Gauss-Jordan with partial pivoting, long double */
bool invert(int n, const std::vector<double> &A, std::vector<double> &Ai)
{
    std::vector<long double> a(n * 2 * n, 0.0L);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) a[i * 2 * n + j] = A[i * n + j];
        a[i * 2 * n + n + i] = 1.0L;
    }
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs((double)a[r * 2 * n + c]) > std::fabs((double)a[piv * 2 * n + c])) piv = r;
        if (a[piv * 2 * n + c] == 0.0L) return false;
        if (piv != c)
            for (int j = 0; j < 2 * n; j++) std::swap(a[c * 2 * n + j], a[piv * 2 * n + j]);
        long double d = a[c * 2 * n + c];
        for (int j = 0; j < 2 * n; j++) a[c * 2 * n + j] /= d;
        for (int r = 0; r < n; r++)
            if (r != c) {
                long double f = a[r * 2 * n + c];
                if (f != 0.0L)
                    for (int j = 0; j < 2 * n; j++) a[r * 2 * n + j] -= f * a[c * 2 * n + j];
            }
    }
    Ai.assign(n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Ai[i * n + j] = (double)a[i * 2 * n + n + j];
    return true;
}

/* ------------------------------------------------------------------ */
/* point physics                                                        */
/* ------------------------------------------------------------------ */

/* BasicOperations.cpp:6-12, eps = 1e-4 (BasicOperations.hpp:9) */
template <class T> T logmean(T x, T y)
{
    T xi = y / x;
    T u = (xi * (xi - 2.0) + 1.0) / (xi * (xi + 2.0) + 1.0);
    return (u < T(1e-4)) ? (x + y) * 52.5 / (105.0 + u * (35.0 + u * (21.0 + 15.0 * u)))
                         : (y - x) / mlog(xi);
}

/* BasicOperations.cpp:53-56 with b = 0 */
template <class T> T sound_speed(T p, T rho, double gamma) { return msqrt(gamma * p / rho); }

/* MFEM Vector::Norml2 (scaled, overflow-safe) used at ChandrashekarFlux.cpp:88 */
template <class T> T norml2(int n, const T *d)
{
    if (n == 1) return mabs(d[0]);
    T scale = 0.0, sum = 0.0;
    for (int i = 0; i < n; i++) {
        if (d[i] != T(0.0)) {
            T a = mabs(d[i]);
            if (scale <= a) {
                T s = scale / a;
                sum = 1.0 + sum * (s * s);
                scale = a;
                continue;
            }
            T s = a / scale;
            sum += s * s;
        }
    }
    return scale * msqrt(sum);
}

/* ChandrashekarFlux::ComputeVolumeFlux, ChandrashekarFlux.cpp:11-83 */
template <class T>
T volume_flux(int dim, double gamma, const T *s1, const T *s2, const T *m1, const T *m2, T *F)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    const int nv = dim + 2;
    T metric[3], mom[3];
    for (int d = 0; d < dim; d++) metric[d] = (m1[d] + m2[d]) * 0.5;
    T rho1 = s1[0], rho2 = s2[0];
    T rho_ln = logmean(rho1, rho2);
    T Vsq1 = 0.0, Vsq2 = 0.0, Vn = 0.0, h_hat = 0.0;
    for (int d = 0; d < dim; d++) {
        T a1 = s1[1 + d] / rho1, a2 = s2[1 + d] / rho2;
        T am = 0.5 * (a1 + a2);
        if (d == 0) {
            Vsq1 = a1 * a1; Vsq2 = a2 * a2; Vn = am * metric[0];
            h_hat = -0.25 * (a1 * a1 + a2 * a2) + am * am;
        } else {
            Vsq1 += a1 * a1; Vsq2 += a2 * a2; Vn += am * metric[d];
            h_hat += -0.25 * (a1 * a1 + a2 * a2) + am * am;
        }
        mom[d] = rho_ln * am;
    }
    T p1 = gm1 * (s1[nv - 1] - 0.5 * rho1 * Vsq1);
    T p2 = gm1 * (s2[nv - 1] - 0.5 * rho2 * Vsq2);
    T lam = msqrt(Vsq1) + sound_speed(p1, rho1, gamma);
    lam = mmax(lam, msqrt(Vsq2) + sound_speed(p2, rho2, gamma));
    T beta1 = 0.5 * rho1 / p1, beta2 = 0.5 * rho2 / p2;
    T beta_ln = logmean(beta1, beta2);
    T p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);
    h_hat += 0.5 / beta_ln * gm1i + p_hat / rho_ln;
    F[0] = rho_ln * Vn;
    for (int d = 0; d < dim; d++) F[1 + d] = Vn * mom[d] + p_hat * metric[d];
    F[1 + dim] = rho_ln * Vn * h_hat;
    return lam;
}

/* ChandrashekarFlux::ComputeFaceFlux, ChandrashekarFlux.cpp:85-161.
 * quirks=1 keeps the 3-D energy-dissipation term of line 131
 * ("+ w_mean" where the 2-D line 118 multiplies). */
template <class T>
T face_flux(int dim, double gamma, int quirks, const T *s1, const T *s2, const T *nor, T *F)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    const int nv = dim + 2;
    T nor_mag = norml2(dim, nor);
    T rho1 = s1[0], rho2 = s2[0];
    T rho_mean = 0.5 * (rho1 + rho2);
    T rho_ln = logmean(rho1, rho2);
    T drho = rho2 - rho1;
    T Vsq1 = 0.0, Vsq2 = 0.0, Vn = 0.0, h_hat = 0.0, diss = 0.0, mom[3];
    for (int d = 0; d < dim; d++) {
        T a1 = s1[1 + d] / rho1, a2 = s2[1 + d] / rho2;
        T am = 0.5 * (a1 + a2);
        T da = a2 - a1;
        if (d == 0) {
            Vsq1 = a1 * a1; Vsq2 = a2 * a2; Vn = am * nor[0];
            h_hat = -0.25 * (a1 * a1 + a2 * a2) + am * am;
            diss = 0.5 * drho * a1 * a2 + rho_mean * da * am;
        } else {
            Vsq1 += a1 * a1; Vsq2 += a2 * a2; Vn += am * nor[d];
            h_hat += -0.25 * (a1 * a1 + a2 * a2) + am * am;
            if (d == 2 && quirks)
                diss += 0.5 * drho * a1 * a2 + rho_mean * da + am;
            else
                diss += 0.5 * drho * a1 * a2 + rho_mean * da * am;
        }
        mom[d] = rho_ln * am;
    }
    T p1 = gm1 * (s1[nv - 1] - 0.5 * rho1 * Vsq1);
    T p2 = gm1 * (s2[nv - 1] - 0.5 * rho2 * Vsq2);
    T lam = msqrt(Vsq1) + sound_speed(p1, rho1, gamma);
    lam = mmax(lam, msqrt(Vsq2) + sound_speed(p2, rho2, gamma));
    T beta1 = 0.5 * rho1 / p1, beta2 = 0.5 * rho2 / p2;
    T beta_ln = logmean(beta1, beta2);
    T p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);
    h_hat += 0.5 / beta_ln * gm1i + p_hat / rho_ln;
    diss += 0.5 * drho * gm1i / beta_ln + 0.5 * rho_mean * gm1i * (1.0 / beta2 - 1.0 / beta1);
    F[0] = rho_ln * Vn - 0.5 * lam * (rho2 - rho1) * nor_mag;
    for (int d = 0; d < dim; d++)
        F[1 + d] = Vn * mom[d] + p_hat * nor[d] - 0.5 * lam * (s2[1 + d] - s1[1 + d]) * nor_mag;
    F[1 + dim] = rho_ln * Vn * h_hat - 0.5 * lam * diss * nor_mag;
    return lam;
}

/* Normalize (BasicOperations.hpp:39-42): vec /= Norml2(vec), MFEM's
 * operator/= multiplying by the reciprocal */
template <class T> void normalize(int dim, const T *nor, T *unit)
{
    T m = 1.0 / norml2(dim, nor);
    for (int d = 0; d < dim; d++) unit[d] = nor[d] * m;
}

/* MFEM Vector dot product: prod = 0, prod += a_i b_i in index order */
template <class T> T dot(int n, const T *a, const T *b)
{
    T prod = 0.0;
    for (int i = 0; i < n; i++) prod += a[i] * b[i];
    return prod;
}

/* RotateState (BasicOperations.cpp:238-259) with Normal and Cross
 * (BasicOperations.cpp:211-236): momentum components in the (n, t1, t2)
 * frame, t1 = Normal(n), t2 = n x t1 */
template <class T> void rotate_state(int dim, T *state, const T *nor)
{
    T tan1[3];
    if (dim == 2) { tan1[0] = -nor[1]; tan1[1] = nor[0]; }
    else {
        if (mabs(nor[0]) > T(1e-10)) {
            T q = nor[1] / nor[0];
            T den = msqrt(1.0 + mpow(q, 2.0));
            tan1[0] = -(nor[1] / nor[0]) / den;
            tan1[1] = 1.0 / den;
        } else {
            tan1[0] = -1.0;
            tan1[1] = 0.0;
        }
        tan1[2] = 0.0;
    }
    T *rhoV = state + 1;
    T rho_u = dot(dim, rhoV, nor);
    T rho_v = dot(dim, rhoV, tan1);
    if (dim == 3) {
        T tan2[3];
        tan2[0] = nor[1] * tan1[2] - nor[2] * tan1[1];
        tan2[1] = nor[2] * tan1[0] - nor[0] * tan1[2];
        tan2[2] = nor[0] * tan1[1] - nor[1] * tan1[0];
        rhoV[2] = dot(dim, rhoV, tan2);
    }
    rhoV[0] = rho_u;
    rhoV[1] = rho_v;
}

/* SlipWallBdrFaceIntegrator::ComputeBdrFaceInviscidFlux
 * (SlipWallBdrFaceIntegrator.cpp:15-60): the wall pressure p* of the
 * state rotated into the face frame (shock branch for u_n > 0, rarefaction
 * branch clipped at 1e-4 otherwise), flux (0, p* nor, 0); returns the
 * rotated-frame |v| + a. */
template <class T> T slip_wall_flux(int dim, double gamma, const T *s1, const T *nor, T *F)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    const double gp1 = 1.0 + gamma, gp1i = 1.0 / gp1;
    const int nv = dim + 2;
    T unit[3], st[5], prim[5];
    normalize(dim, nor, unit);
    for (int v = 0; v < nv; v++) st[v] = s1[v];
    if (dim == 1) st[1] = st[1] * nor[0];   /* SlipWallBdrFaceIntegrator.cpp:28-30 */
    else rotate_state(dim, st, unit);
    /* Conserv2Prim (BasicOperations.cpp:172-186): velocities by division,
     * pressure by ComputePressure */
    prim[0] = st[0];
    for (int d = 0; d < dim; d++) prim[1 + d] = st[1 + d] / st[0];
    {
        T r = 1.0 / st[0];
        T V[3];
        for (int d = 0; d < dim; d++) V[d] = st[1 + d] * r;
        prim[nv - 1] = gm1 * (st[nv - 1] - 0.5 * st[0] * dot(dim, V, V));
    }
    T rho = prim[0], un = prim[1], p = prim[nv - 1];
    T an = sound_speed(p, rho, gamma);
    T p_star;
    if (un > T(0.0))
        p_star = p + 0.25 * un * gp1 * rho *
                 (un + msqrt(mpow(un, 2.0) + 8.0 * gp1i / rho * p * (gm1 * gp1i + 1.0)));
    else
        p_star = p * mpow(mmax(1.0 + 0.5 * gm1 * un / an, T(0.0001)), 2.0 * gamma * gm1i);
    for (int v = 0; v < nv; v++) F[v] = 0.0;
    for (int d = 0; d < dim; d++) F[1 + d] = p_star * nor[d];
    return msqrt(dot(dim, prim + 1, prim + 1)) + an;
}

/* ComputeBdrFaceInviscidFlux of each boundary type: the outer state of
 * ComputeOuterInviscidState and ChandrashekarFlux::ComputeFaceFlux with
 * state1 = interior (BdrFaceIntegrator.cpp:252-258); slip walls replace the
 * whole flux.  Symmetry (SymmetryBdrFaceIntegrator.cpp): outer momentum
 * m - 2 (m.n) n with the unit normal, a zero normal gives a zero flux. */
template <class T>
T bc_flux(int dim, double gamma, int quirks, int type, const T *s1, const T *outer, const T *nor, T *F)
{
    const int nv = dim + 2;
    T s2[5];
    for (int v = 0; v < nv; v++) s2[v] = s1[v];
    switch (type) {
    case ORC_BC_SLIP:
    case ORC_BC_NO_SLIP_ADIABATIC:   /* inherits SlipWall's inviscid flux */
        return slip_wall_flux(dim, gamma, s1, nor, F);
    case ORC_BC_SUPERSONIC_INFLOW:
    case ORC_BC_SPECIFIED_STATE:
        for (int v = 0; v < nv; v++) s2[v] = outer[v];
        break;
    case ORC_BC_SUPERSONIC_OUTFLOW:
        for (int v = 0; v < nv; v++) s2[v] = s1[v];
        break;
    case ORC_BC_SYMMETRY: {
        if (!(norml2(dim, nor) != T(0.0))) {
            for (int v = 0; v < nv; v++) F[v] = 0.0;
            return 0.0;
        }
        T unit[3];
        normalize(dim, nor, unit);
        for (int v = 0; v < nv; v++) s2[v] = s1[v];
        T mn = dot(dim, s2 + 1, unit);
        for (int d = 0; d < dim; d++) s2[1 + d] += -2.0 * mn * unit[d];
        break;
    }
    default:
        for (int v = 0; v < nv; v++) F[v] = 0.0;
        return 0.0;
    }
    return face_flux(dim, gamma, quirks, s1, s2, nor, F);
}

/* Entropy2Conserv, BasicOperations.cpp:106-122.  `Vector vel(ent_state.
 * GetData() + 1, dim); vel /= beta` scales the caller's entropy vector in
 * place (operator/= multiplies by 1/beta): on return w[1..dim] hold the
 * velocities, which the faithful boundary lifting flux then subtracts. */
template <class T> void entropy2conserv(int dim, double gamma, T *w, T *s)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    T beta = -w[dim + 1];
    T ib = 1.0 / beta;
    for (int d = 0; d < dim; d++) w[1 + d] = w[1 + d] * ib;
    T vv = 0.0;
    for (int d = 0; d < dim; d++) vv += w[1 + d] * w[1 + d];
    T sent = gamma - (w[0] + 0.5 * beta * vv) * gm1;
    s[0] = mpow(mexp(-sent) / beta, gm1i);
    for (int d = 0; d < dim; d++) s[1 + d] = s[0] * w[1 + d];
    s[dim + 1] = s[0] * (1.0 / beta * gm1i + 0.5 * vv);
}

template <class T> void cons2entropy(int dim, double gamma, const T *s, T *w);

/* ComputeBdrFaceLiftingFlux.  Base (BdrFaceIntegrator.cpp:276-282):
 * 0.5 (w(u_out) - w), u_out the outer state of the conservative interior
 * state; quirks=1 keeps the velocities Entropy2Conserv leaves in w.  No-slip
 * adiabatic wall (NoSlipAdiabWallBdrFaceIntegrator.cpp): the wall's entropy
 * variables (w0, beta V, -beta) minus w, beta = -w_last, not halved. */
template <class T>
void bc_lift_flux(int dim, double gamma, int quirks, int type, const T *w, const T *outer, T *f)
{
    const int nv = dim + 2;
    if (type == ORC_BC_NO_SLIP_ADIABATIC) {
        T v = -w[nv - 1];
        f[0] = w[0];
        for (int d = 0; d < dim; d++) f[1 + d] = T(outer[d]) * v;
        f[nv - 1] = -v;
        for (int k = 0; k < nv; k++) f[k] -= w[k];
        return;
    }
    T wi[5], cs[5], g[5];
    for (int k = 0; k < nv; k++) wi[k] = w[k];
    entropy2conserv(dim, gamma, wi, cs);
    const T *ghost = (type == ORC_BC_SUPERSONIC_OUTFLOW) ? cs : outer;
    cons2entropy(dim, gamma, ghost, g);
    for (int k = 0; k < nv; k++) f[k] = (g[k] - (quirks ? wi[k] : w[k])) * 0.5;
}

/* ComputePressure, BasicOperations.cpp:31-37 (V /= rho multiplies by 1/rho) */
template <class T> T pressure(int dim, double gamma, const T *s)
{
    T r = 1.0 / s[0];
    T vv = 0.0;
    for (int d = 0; d < dim; d++) {
        T v = s[1 + d] * r;
        vv += v * v;
    }
    return (gamma - 1.0) * (s[dim + 1] - 0.5 * s[0] * vv);
}

/* Conserv2Entropy, BasicOperations.cpp:81-104 (beta = rho/p here) */
template <class T> void cons2entropy(int dim, double gamma, const T *s, T *w)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    T r = 1.0 / s[0];
    T vel[3];
    T vv = 0.0;
    for (int d = 0; d < dim; d++) { vel[d] = s[1 + d] * r; vv += vel[d] * vel[d]; }
    T v_sq = 0.5 * vv;
    T p = gm1 * (s[dim + 1] - s[0] * v_sq);
    T beta = s[0] / p;
    T sent = mlog(p * mpow(s[0], -gamma));
    w[0] = (gamma - sent) * gm1i - beta * v_sq;
    for (int d = 0; d < dim; d++) w[1 + d] = beta * vel[d];
    w[dim + 1] = -beta;
}

/* EntropyGrad2PrimGrad, BasicOperations.cpp:126-170, one direction, in place.
 * quirks=1: `Vector vel(state.GetData()+1, dim); vel /= state(0)` (lines
 * 137-138) aliases the row copy, so the density-gradient term at line 164
 * multiplies velocities where momenta are meant.  That term cancels from
 * grad p - (p/rho) grad rho, so the viscous flux is unchanged up to
 * rounding; only grad rho and grad p themselves differ. */
template <class T> void entgrad2primgrad(int dim, double gamma, int quirks, const T *s, T *g)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    T r = 1.0 / s[0];
    T vel[3];
    T vv = 0.0;
    for (int d = 0; d < dim; d++) { vel[d] = s[1 + d] * r; vv += vel[d] * vel[d]; }
    T v_sq = 0.5 * vv;
    T p = gm1 * (s[dim + 1] - s[0] * v_sq);
    T KE = v_sq * s[0];
    T gs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < dim + 2; v++) gs[v] = g[v];
    const int L = dim + 1;
    T mdot = 0.0;
    for (int d = 0; d < dim; d++) {
        g[1 + d] = p / s[0] * (gs[1 + d] + vel[d] * gs[L]);
        T md = quirks ? vel[d] : s[1 + d];
        if (d == 0) mdot = md * g[1];
        else mdot += md * g[1 + d];
    }
    g[0] = s[0] * gs[0] - gs[L] * (KE - p * gm1i) + s[0] / p * mdot;
    g[L] = p / s[0] * (g[0] + p * gs[L]);
}

/* NavierStokesFlux::ComputeViscosity (NavierStokesFlux.hpp:32-36): Sutherland's
 * law with the constants of Physics.hpp:35-37 and T0pTs = T0 + Ts formed as
 * in the constructor (NavierStokesFlux.hpp:22). */
template <class T>
T sutherland_mu(T rho, T p, double R_gas)
{
    const double mu0 = 1.716e-5, T0 = 273.15, Ts = 110.4, T0pTs = T0 + Ts;
    const T temp = p / (rho * T(R_gas));
    return T(mu0) * T(T0pTs) / (temp + T(Ts)) * (temp / T(T0)) * msqrt(temp / T(T0));
}

/* NavierStokesFlux::ComputeViscousFlux, NavierStokesFlux.cpp:12-90 (mu_bulk =
 * 2/3 from Physics.hpp:34).  R_gas > 0: the SUTHERLAND build's per-point mu
 * (lines 34-36), else the constant mu.  flux is nv x dim, row 0 = 0. */
template <class T>
void viscous_flux(int dim, double gamma, double Pr, double mu_const, double R_gas, const T *s,
                  const T *const *dq, T *flux)
{
    const double gm1 = gamma - 1.0, gm1i = 1.0 / gm1;
    const double mu_bulk = 2.0 / 3.0;
    const int nv = dim + 2;
    T prim[5];
    prim[0] = s[0];
    for (int d = 0; d < dim; d++) prim[1 + d] = s[1 + d] / s[0];
    prim[nv - 1] = pressure(dim, gamma, s);
    const T mu = R_gas > 0.0 ? sutherland_mu<T>(prim[0], prim[nv - 1], R_gas) : T(mu_const);
    T lambda = mu * gamma * (1.0 / Pr);
    T div = dq[0][1];
    for (int d = 1; d < dim; d++) div += dq[d][1 + d];
    T cvdT[3];
    for (int c = 0; c < dim; c++)
        cvdT[c] = gm1i / prim[0] * (dq[c][nv - 1] - prim[nv - 1] / prim[0] * dq[c][0]);
    for (int c = 0; c < dim; c++) flux[0 * dim + c] = 0.0;
    for (int c = 0; c < dim; c++) {
        for (int a = 0; a < dim; a++) {
            /* tau(a, c): mu (2 du_a/dx_a - mu_bulk div) on the diagonal,
             * mu (du_a/dx_c + du_c/dx_a) off it, with the reference's
             * operand order (first index is the row of the flux matrix) */
            if (a == c) flux[(1 + a) * dim + c] = mu * (2.0 * dq[a][1 + a] - mu_bulk * div);
            else if (c == 0) flux[(1 + a) * dim + c] = mu * (dq[a][1 + 0] + dq[0][1 + a]);
            else flux[(1 + a) * dim + c] = mu * (dq[c][1 + a] + dq[a][1 + c]);
        }
        T e = prim[1] * flux[1 * dim + c];
        for (int a = 1; a < dim; a++) e = e + prim[1 + a] * flux[(1 + a) * dim + c];
        flux[(nv - 1) * dim + c] = e + lambda * cvdT[c];
    }
}

/* ------------------------------------------------------------------ */
/* mesh                                                                 */
/* ------------------------------------------------------------------ */
struct Mesh {
    int dim, N, Np, K, NE, Nfp, nv;
    int n[3];
    double h[3], lo[3];
    std::vector<double> J, adj;      /* [NE*K], [NE*K*dim*dim] row r = Ja^r */
    int nf;
    std::vector<int> fe1, fe2, fn1, fn2;
    std::vector<double> fnor;        /* [nf*Nfp*dim] */
    std::vector<double> sub;         /* [NE][dim][Nfp][N][dim] interior subcell normals */
    double hmin;
    /* boundary faces: elements at the low (side 0) / high (side 1) end of
     * each non-periodic direction r; their faces' elem1 is that element */
    std::vector<int> bel[3][2];
    double ad[3];                    /* face measure of an r-face */
    /* quadrilateral meshes: node coordinates, and the boundary faces
     * (element, type, attribute, face nodes and outward normals per point) */
    bool quad = false;
    std::vector<double> X;           /* [2][NE*K] */
    std::vector<int> bq_e, bq_type, bq_attr, bq_n;
    std::vector<double> bq_nor;
    std::vector<double> attr_state; /* [nattr][5] */
    /* axisymmetric quadrilateral meshes: element corners [NE][4][2], the
     * nodes in axis_idx (BuildAxisIndexFromMarker, DGSEMOperator.cpp:
     * 446-485) and whether an axis boundary exists */
    std::vector<double> cq;
    std::vector<char> axis_idx;
    bool axis_bc = false;
};

inline int line_node(int dim, int Np, int r, int l, int m)
{
    if (dim == 1) return m;
    if (dim == 2) return r == 0 ? l * Np + m : m * Np + l;
    int a = l % Np, b = l / Np;
    if (r == 0) return (b * Np + a) * Np + m;
    if (r == 1) return (b * Np + m) * Np + a;
    return (m * Np + b) * Np + a;
}

} /* namespace */

struct orc_ctx {
    orc_params p;
    Mesh m;
    std::vector<double> x, w, D, Dh, Dh2;
    std::vector<double> Vinv;        /* K x K */
    std::vector<int> degs;           /* K x dim (ModalBasis ubdegs) */
    std::vector<double> alpha, eta, wJ;
    std::vector<double> grad;        /* [dim][nv][NE][K] */
    std::vector<double> stage_y, stage_k;
    double lam;
    double thresh;
    /* time of the next RHS (TimeDependentOperator::t) and of the RHS being
     * evaluated; per boundary attribute of a quadrilateral mesh an optional
     * split outer state [kind, n0, n1, n2, offset, speed, below[5], above[5]] */
    double time = 0.0, t_stage = 0.0;
    std::vector<double> attr_fn;
};

namespace {

void build_subcells(orc_ctx *c);

void build_mesh(orc_ctx *c)
{
    Mesh &m = c->m;
    const orc_params &p = c->p;
    m.dim = p.dim; m.N = p.order; m.Np = p.order + 1;
    m.nv = p.dim + 2;
    m.K = m.Np * (p.dim >= 2 ? m.Np : 1) * (p.dim == 3 ? m.Np : 1);
    m.Nfp = (p.dim == 3) ? m.Np * m.Np : p.dim == 2 ? m.Np : 1;
    m.NE = 1;
    for (int r = 0; r < 3; r++) m.n[r] = (r < p.dim) ? p.nelem[r] : 1;
    for (int r = 0; r < p.dim; r++) {
        m.NE *= m.n[r];
        m.h[r] = (p.hi[r] - p.lo[r]) / m.n[r];
        m.lo[r] = p.lo[r];
    }
    const int dim = m.dim, K = m.K, NE = m.NE, Np = m.Np, N = m.N, Nfp = m.Nfp;
    /* Cartesian Jacobian diag(h); det and adjugate with MFEM's association */
    double Jd, ad[3] = {0, 0, 0};
    /* 1-D: MFEM's adjugate of a 1 x 1 matrix is 1, so the face normal is
     * +-1 (DGSEMIntegrator.cpp:180-183 set it from the face's side) */
    if (dim == 1) { Jd = m.h[0]; ad[0] = 1.0; }
    else if (dim == 2) { Jd = m.h[0] * m.h[1]; ad[0] = m.h[1]; ad[1] = m.h[0]; }
    else {
        Jd = m.h[0] * (m.h[1] * m.h[2]);
        ad[0] = m.h[1] * m.h[2]; ad[1] = m.h[0] * m.h[2]; ad[2] = m.h[0] * m.h[1];
    }
    m.hmin = m.h[0];
    for (int r = 1; r < dim; r++) m.hmin = std::min(m.hmin, m.h[r]);
    m.J.assign((size_t)NE * K, Jd);
    m.adj.assign((size_t)NE * K * dim * dim, 0.0);
    for (size_t q = 0; q < (size_t)NE * K; q++)
        for (int r = 0; r < dim; r++) m.adj[q * dim * dim + r * dim + r] = ad[r];
    for (int r = 0; r < 3; r++) m.ad[r] = ad[r];
    /* interior faces: every element is elem1 of its +r face, elem2 is its
     * periodic +r neighbour, normal = +|face| e_r; in a non-periodic
     * direction the +r face of the last element is a boundary face
     * (fe2 = -1, skipped by the interior loops) */
    m.nf = dim * NE;
    m.fe1.resize(m.nf); m.fe2.resize(m.nf);
    m.fn1.resize((size_t)m.nf * Nfp); m.fn2.resize((size_t)m.nf * Nfp);
    m.fnor.assign((size_t)m.nf * Nfp * dim, 0.0);
    int f = 0;
    for (int r = 0; r < dim; r++)
        for (int e = 0; e < NE; e++, f++) {
            int ex = e % m.n[0], ey = (e / m.n[0]) % m.n[1], ez = e / (m.n[0] * m.n[1]);
            int idx[3] = {ex, ey, ez};
            const bool open = p.bc[r][0] != ORC_BC_PERIODIC;
            if (open && idx[r] == 0) m.bel[r][0].push_back(e);
            if (open && idx[r] == m.n[r] - 1) m.bel[r][1].push_back(e);
            const bool last = idx[r] + 1 == m.n[r];
            idx[r] = (idx[r] + 1) % m.n[r];
            int e2 = (idx[2] * m.n[1] + idx[1]) * m.n[0] + idx[0];
            m.fe1[f] = e; m.fe2[f] = (open && last) ? -1 : e2;
            for (int q = 0; q < Nfp; q++) {
                m.fn1[(size_t)f * Nfp + q] = line_node(dim, Np, r, q, N);
                m.fn2[(size_t)f * Nfp + q] = line_node(dim, Np, r, q, 0);
                m.fnor[((size_t)f * Nfp + q) * dim + r] = ad[r];
            }
        }
    build_subcells(c);
}

/* subcell normals, DGSEMIntegrator.cpp:1199-1304:
 * n_{m+1/2} = Ja^r(line start) + sum_{l<=m} w_l sum_q D_{lq} Ja^r_q */
void build_subcells(orc_ctx *c)
{
    Mesh &m = c->m;
    const int dim = m.dim, K = m.K, NE = m.NE, Np = m.Np, N = m.N, Nfp = m.Nfp;
    if (c->p.blending) {
        m.sub.assign((size_t)NE * dim * Nfp * N * dim, 0.0);
        for (int e = 0; e < NE; e++)
            for (int r = 0; r < dim; r++)
                for (int l = 0; l < Nfp; l++) {
                    int n0 = line_node(dim, Np, r, l, 0);
                    const double *m1 = &m.adj[((size_t)e * K + n0) * dim * dim + r * dim];
                    for (int i = 1; i < Np; i++) {
                        double nor[3];
                        for (int d = 0; d < dim; d++) nor[d] = m1[d];
                        for (int ll = 0; ll < i; ll++) {
                            double tmp[3] = {0, 0, 0};
                            for (int q = 0; q < Np; q++) {
                                int nq = line_node(dim, Np, r, l, q);
                                const double *m2 = &m.adj[((size_t)e * K + nq) * dim * dim + r * dim];
                                for (int d = 0; d < dim; d++) tmp[d] += m2[d] * c->D[ll * Np + q];
                            }
                            for (int d = 0; d < dim; d++) { tmp[d] *= c->w[ll]; nor[d] += tmp[d]; }
                        }
                        double *dst = &m.sub[((((size_t)e * dim + r) * Nfp + l) * N + (i - 1)) * dim];
                        for (int d = 0; d < dim; d++) dst[d] = nor[d];
                    }
                }
    }
}

/* Quadrilateral mesh geometry, restating what mfem::Mesh and its element
 * transformations give the reference (DGSEMIntegrator.cpp:251-267, 186;
 * BdrFaceIntegrator.cpp:79-95):
 *  - node (i, j) of element e sits at (xi_i, eta_j) of the bilinear map of
 *    its four corners; the Jacobian there is the sum over corners of corner
 *    times the bilinear shape derivative (BiLinear2DFiniteElement), its
 *    adjugate rows are the contravariant metrics, its determinant J;
 *  - an edge shared by two elements is one face, its first element the one
 *    with the lower index (faces are created in element order); the normal
 *    at every point is CalcOrtho of the first element's edge Jacobian,
 *    (dy, -dx) along its counterclockwise traversal, outward of it;
 *  - an edge that no other element shares is a boundary face of the
 *    segment with the same two vertices.
 * Returns 0, or 8 for a mesh that does not close. */
int build_mesh_quad(orc_ctx *c, const orc_quad_mesh *q)
{
    Mesh &m = c->m;
    m.quad = true;
    m.dim = 2; m.N = c->p.order; m.Np = m.N + 1; m.nv = 4;
    m.K = m.Np * m.Np; m.Nfp = m.Np;
    m.NE = q->nelem;
    m.n[0] = m.NE; m.n[1] = m.n[2] = 1;
    const int K = m.K, NE = m.NE, Np = m.Np, N = m.N, Nfp = m.Nfp;
    const size_t NK = (size_t)NE * K;
    auto cornerx = [&](int e, int k, int i) {
        return q->elem_coords ? q->elem_coords[(size_t)e * 8 + 2 * k + i]
                              : q->vert[2 * (size_t)q->elem[(size_t)e * 4 + k] + i];
    };
    m.J.assign(NK, 0.0);
    m.adj.assign(NK * 4, 0.0);
    m.X.assign(2 * NK, 0.0);
    m.cq.assign((size_t)NE * 8, 0.0);
    for (int e = 0; e < NE; e++)
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < 2; i++) m.cq[(size_t)e * 8 + 2 * k + i] = cornerx(e, k, i);
    m.hmin = 1e300;
    for (int e = 0; e < NE; e++) {
        for (int j = 0; j < Np; j++)
            for (int i = 0; i < Np; i++) {
                const double xi = c->x[i], eta = c->x[j];
                /* shape N_k and derivatives (dN/dxi, dN/deta), corners in vertex order */
                const double Nk[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta};
                const double dN[4][2] = {{eta - 1, xi - 1}, {1 - eta, -xi}, {eta, xi}, {-eta, 1 - xi}};
                double Jm[2][2];
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) {
                        double sum = 0.0;
                        for (int k = 0; k < 4; k++) sum += dN[k][b] * cornerx(e, k, a);
                        Jm[a][b] = sum;
                    }
                const size_t nd = (size_t)e * K + j * Np + i;
                for (int a = 0; a < 2; a++) {
                    double x = 0.0;
                    for (int k = 0; k < 4; k++) x += Nk[k] * cornerx(e, k, a);
                    m.X[a * NK + nd] = x;
                }
                /* adjugate: row 0 = (J11, -J01), row 1 = (-J10, J00) */
                double *ad = &m.adj[nd * 4];
                ad[0] = Jm[1][1]; ad[1] = -Jm[0][1];
                ad[2] = -Jm[1][0]; ad[3] = Jm[0][0];
                m.J[nd] = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
                if (!(m.J[nd] > 0.0)) return 8;
            }
        /* GetElementSize(e, 1): smaller singular value of J at the centre */
        double Jc[2][2];
        const double dN[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                double sum = 0.0;
                for (int k = 0; k < 4; k++) sum += dN[k][b] * cornerx(e, k, a);
                Jc[a][b] = sum;
            }
        const double S = Jc[0][0] * Jc[0][0] + Jc[0][1] * Jc[0][1] + Jc[1][0] * Jc[1][0] + Jc[1][1] * Jc[1][1];
        const double det = std::fabs(Jc[0][0] * Jc[1][1] - Jc[1][0] * Jc[0][1]);
        const double smin = std::sqrt(0.5 * (S - std::sqrt(std::max(S * S - 4 * det * det, 0.0))));
        m.hmin = std::min(m.hmin, smin);
    }
    /* edges of local face f = 2r + side in increasing-l order, and the
     * counterclockwise traversal of the same edge */
    const int fv[4][2] = {{0, 3}, {1, 2}, {0, 1}, {3, 2}};
    const int ccw[4][2] = {{3, 0}, {1, 2}, {0, 1}, {2, 3}};
    auto vtx = [&](int e, int k) { return q->elem[(size_t)e * 4 + k]; };
    std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> edges;
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 4; f++) {
            int a = vtx(e, fv[f][0]), b = vtx(e, fv[f][1]);
            edges[{std::min(a, b), std::max(a, b)}].push_back({e, f});
        }
    std::map<std::pair<int, int>, int> seg;
    for (int i = 0; i < q->nbdr; i++) {
        int a = q->bdr[2 * i], b = q->bdr[2 * i + 1];
        seg[{std::min(a, b), std::max(a, b)}] = q->bdr_attr[i];
    }
    m.fe1.clear(); m.fe2.clear(); m.fn1.clear(); m.fn2.clear(); m.fnor.clear();
    for (auto &kv : edges) {
        const auto &lst = kv.second;
        if (lst.size() > 2) return 8;
        int e1 = lst[0].first, f1 = lst[0].second;
        if (lst.size() == 2 && lst[1].first < e1) { e1 = lst[1].first; f1 = lst[1].second; }
        const double px = cornerx(e1, ccw[f1][0], 0), py = cornerx(e1, ccw[f1][0], 1);
        const double qx = cornerx(e1, ccw[f1][1], 0), qy = cornerx(e1, ccw[f1][1], 1);
        const double nx = qy - py, ny = -(qx - px);
        const int r1 = f1 >> 1, s1 = f1 & 1;
        if (lst.size() == 2) {
            const int e2 = (e1 == lst[0].first && f1 == lst[0].second) ? lst[1].first : lst[0].first;
            const int f2 = (e1 == lst[0].first && f1 == lst[0].second) ? lst[1].second : lst[0].second;
            const int r2 = f2 >> 1, s2 = f2 & 1;
            const bool same = vtx(e1, fv[f1][0]) == vtx(e2, fv[f2][0]);
            m.fe1.push_back(e1);
            m.fe2.push_back(e2);
            for (int l = 0; l < Nfp; l++) {
                m.fn1.push_back(line_node(2, Np, r1, l, s1 ? N : 0));
                m.fn2.push_back(line_node(2, Np, r2, same ? l : N - l, s2 ? N : 0));
                m.fnor.push_back(nx);
                m.fnor.push_back(ny);
            }
        } else {
            auto it = seg.find(kv.first);
            if (it == seg.end()) return 8;
            const int attr = it->second;
            if (attr < 1 || attr > q->nattr) return 8;
            const int type = q->attr_bc[attr - 1];
            if (type < ORC_BC_SLIP || type > ORC_BC_AXIS) return 8;
            m.bq_e.push_back(e1);
            /* "axis" adds the symmetry integrator (Simulation.cpp:520-524) */
            m.bq_type.push_back(type == ORC_BC_AXIS ? ORC_BC_SYMMETRY : type);
            if (type == ORC_BC_AXIS) {
                /* every node of the adjacent element with |r| <= 1e-14 */
                m.axis_bc = true;
                if (m.axis_idx.empty()) m.axis_idx.assign(NK, 0);
                for (int i = 0; i < K; i++)
                    if (std::fabs(m.X[NK + (size_t)e1 * K + i]) <= 1e-14) m.axis_idx[(size_t)e1 * K + i] = 1;
            }
            m.bq_attr.push_back(attr - 1);
            for (int l = 0; l < Nfp; l++) {
                m.bq_n.push_back(line_node(2, Np, r1, l, s1 ? N : 0));
                m.bq_nor.push_back(nx);
                m.bq_nor.push_back(ny);
            }
        }
    }
    m.nf = (int)m.fe1.size();
    if (m.axis_idx.empty()) m.axis_idx.assign(NK, 0);
    m.attr_state.assign(5 * (size_t)std::max(q->nattr, 1), 0.0);
    for (int a = 0; a < q->nattr; a++)
        if (q->attr_state)
            for (int v = 0; v < 4; v++) m.attr_state[5 * (size_t)a + v] = q->attr_state[4 * (size_t)a + v];
    build_subcells(c);
    return 0;
}

/* Hexahedral mesh geometry, restating what mfem::Mesh and its element
 * transformations give the reference for 3-D meshes (DGSEMIntegrator.cpp:
 * 251-267, 186; BdrFaceIntegrator.cpp:79-95):
 *  - node (i, j, k) of element e sits at (xi_i, eta_j, zeta_k) of the
 *    trilinear map of its eight corners (MFEM's vertex order: the zeta = 0
 *    face counterclockwise, then zeta = 1); the Jacobian there is the sum
 *    over corners of corner times the trilinear shape derivative, its
 *    adjugate rows the contravariant metrics, its determinant J;
 *  - a quadrilateral face shared by two elements is one face, its first
 *    element the lower index; its points pair up through the four shared
 *    vertices (any of the 8 relative orientations); the normal at each
 *    point is CalcOrtho of the face Jacobian, the cross product of the two
 *    tangents, outward of the first element: the first element's adjugate
 *    row r at that node, with the sign of its side;
 *  - a face no other element shares is a boundary face of the boundary
 *    quadrilateral with the same four vertices.
 * Returns 0, or 8 for a mesh that does not close or has J <= 0. */
int build_mesh_hex(orc_ctx *c, const orc_hex_mesh *q)
{
    Mesh &m = c->m;
    m.quad = true;
    m.dim = 3; m.N = c->p.order; m.Np = m.N + 1; m.nv = 5;
    m.K = m.Np * m.Np * m.Np; m.Nfp = m.Np * m.Np;
    m.NE = q->nelem;
    m.n[0] = m.NE; m.n[1] = m.n[2] = 1;
    const int K = m.K, NE = m.NE, Np = m.Np, N = m.N, Nfp = m.Nfp;
    const size_t NK = (size_t)NE * K;
    auto cornerx = [&](int e, int k, int i) {
        return q->elem_coords ? q->elem_coords[(size_t)e * 24 + 3 * k + i]
                              : q->vert[3 * (size_t)q->elem[(size_t)e * 8 + k] + i];
    };
    /* reference-cube corners of MFEM's vertex order */
    const int cv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    auto jac = [&](int e, const double xi[3], double Jm[3][3], double X[3]) {
        for (int a = 0; a < 3; a++) {
            X[a] = 0.0;
            for (int b = 0; b < 3; b++) Jm[a][b] = 0.0;
        }
        for (int k = 0; k < 8; k++) {
            double f[3], df[3];
            for (int d = 0; d < 3; d++) {
                f[d] = cv[k][d] ? xi[d] : 1.0 - xi[d];
                df[d] = cv[k][d] ? 1.0 : -1.0;
            }
            const double Nk = f[0] * f[1] * f[2];
            const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
            for (int a = 0; a < 3; a++) {
                const double ca = cornerx(e, k, a);
                X[a] += Nk * ca;
                for (int b = 0; b < 3; b++) Jm[a][b] += dN[b] * ca;
            }
        }
    };
    m.J.assign(NK, 0.0);
    m.adj.assign(NK * 9, 0.0);
    m.X.assign(3 * NK, 0.0);
    m.hmin = 1e300;
    for (int e = 0; e < NE; e++) {
        for (int kk = 0; kk < Np; kk++)
            for (int j = 0; j < Np; j++)
                for (int i = 0; i < Np; i++) {
                    const double xi[3] = {c->x[i], c->x[j], c->x[kk]};
                    double Jm[3][3], X[3];
                    jac(e, xi, Jm, X);
                    const size_t nd = (size_t)e * K + (kk * Np + j) * Np + i;
                    for (int a = 0; a < 3; a++) m.X[a * NK + nd] = X[a];
                    /* adj(J): row r is the contravariant metric Ja^r */
                    double *ad = &m.adj[nd * 9];
                    ad[0] = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
                    ad[1] = Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2];
                    ad[2] = Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1];
                    ad[3] = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
                    ad[4] = Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0];
                    ad[5] = Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2];
                    ad[6] = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
                    ad[7] = Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1];
                    ad[8] = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
                    m.J[nd] = Jm[0][0] * ad[0] + Jm[0][1] * ad[3] + Jm[0][2] * ad[6];
                    if (!(m.J[nd] > 0.0)) return 8;
                }
        /* GetElementSize(e, 1): the smallest singular value of J at the
         * centre, from the eigenvalues of J^T J by cyclic Jacobi rotations */
        const double xc[3] = {0.5, 0.5, 0.5};
        double Jm[3][3], X[3], A[3][3];
        jac(e, xc, Jm, X);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                A[a][b] = 0.0;
                for (int k = 0; k < 3; k++) A[a][b] += Jm[k][a] * Jm[k][b];
            }
        for (int sweep = 0; sweep < 60; sweep++) {
            const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
            if (off <= 1e-34 * (A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2])) break;
            for (int p = 0; p < 2; p++)
                for (int r = p + 1; r < 3; r++) {
                    if (A[p][r] == 0.0) continue;
                    const double th = 0.5 * (A[r][r] - A[p][p]) / A[p][r];
                    const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int k = 0; k < 3; k++) {   /* A <- G^T A G */
                        const double akp = A[k][p], akr = A[k][r];
                        A[k][p] = cs * akp - sn * akr;
                        A[k][r] = sn * akp + cs * akr;
                    }
                    for (int k = 0; k < 3; k++) {
                        const double apk = A[p][k], ark = A[r][k];
                        A[p][k] = cs * apk - sn * ark;
                        A[r][k] = sn * apk + cs * ark;
                    }
                }
        }
        const double lmin = std::min(A[0][0], std::min(A[1][1], A[2][2]));
        m.hmin = std::min(m.hmin, std::sqrt(std::max(lmin, 0.0)));
    }
    /* local face f = 2r + side: its corners at face coordinates (a, b) in
     * the order (0,0), (1,0), (0,1), (1,1), a and b the other two directions
     * in increasing order (face point l = a + Np b) */
    const int fc[6][4] = {{0, 3, 4, 7}, {1, 2, 5, 6}, {0, 1, 4, 5}, {3, 2, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};
    auto vtx = [&](int e, int k) { return q->elem[(size_t)e * 8 + k]; };
    auto key = [&](int a, int b, int cc, int d) {
        std::array<int, 4> k4 = {a, b, cc, d};
        std::sort(k4.begin(), k4.end());
        return k4;
    };
    std::map<std::array<int, 4>, std::vector<std::pair<int, int>>> faces;
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 6; f++)
            faces[key(vtx(e, fc[f][0]), vtx(e, fc[f][1]), vtx(e, fc[f][2]), vtx(e, fc[f][3]))].push_back({e, f});
    std::map<std::array<int, 4>, int> bq;
    for (int i = 0; i < q->nbdr; i++) {
        const int *b = &q->bdr[4 * (size_t)i];
        bq[key(b[0], b[1], b[2], b[3])] = q->bdr_attr[i];
    }
    m.fe1.clear(); m.fe2.clear(); m.fn1.clear(); m.fn2.clear(); m.fnor.clear();
    auto normal = [&](int e, int f, int node, double *nor) {
        const int r = f >> 1;
        const double sg = (f & 1) ? 1.0 : -1.0;
        for (int d = 0; d < 3; d++) nor[d] = sg * m.adj[((size_t)e * K + node) * 9 + r * 3 + d];
    };
    for (auto &kv : faces) {
        const auto &lst = kv.second;
        if (lst.size() > 2) return 8;
        int e1 = lst[0].first, f1 = lst[0].second;
        if (lst.size() == 2 && lst[1].first < e1) { e1 = lst[1].first; f1 = lst[1].second; }
        const int r1 = f1 >> 1, s1 = f1 & 1;
        if (lst.size() == 2) {
            const bool first0 = (e1 == lst[0].first && f1 == lst[0].second);
            const int e2 = first0 ? lst[1].first : lst[0].first;
            const int f2 = first0 ? lst[1].second : lst[0].second;
            const int r2 = f2 >> 1, s2 = f2 & 1;
            /* where the first element's corners (0,0), (1,0), (0,1) sit on
             * the second element's face */
            int ca[3], cb[3];
            const int pick[3] = {0, 1, 2};
            for (int t = 0; t < 3; t++) {
                const int g = vtx(e1, fc[f1][pick[t]]);
                int hit = -1;
                for (int u = 0; u < 4; u++)
                    if (vtx(e2, fc[f2][u]) == g) hit = u;
                if (hit < 0) return 8;
                ca[t] = hit & 1;
                cb[t] = hit >> 1;
            }
            m.fe1.push_back(e1);
            m.fe2.push_back(e2);
            for (int l = 0; l < Nfp; l++) {
                const int a = l % Np, b = l / Np;
                const int a2 = N * ca[0] + (ca[1] - ca[0]) * a + (ca[2] - ca[0]) * b;
                const int b2 = N * cb[0] + (cb[1] - cb[0]) * a + (cb[2] - cb[0]) * b;
                const int n1 = line_node(3, Np, r1, l, s1 ? N : 0);
                m.fn1.push_back(n1);
                m.fn2.push_back(line_node(3, Np, r2, a2 + Np * b2, s2 ? N : 0));
                double nor[3];
                normal(e1, f1, n1, nor);
                for (int d = 0; d < 3; d++) m.fnor.push_back(nor[d]);
            }
        } else {
            auto it = bq.find(kv.first);
            if (it == bq.end()) return 8;
            const int attr = it->second;
            if (attr < 1 || attr > q->nattr) return 8;
            const int type = q->attr_bc[attr - 1];
            if (type < ORC_BC_SLIP || type > ORC_BC_NO_SLIP_ADIABATIC) return 8;
            m.bq_e.push_back(e1);
            m.bq_type.push_back(type);
            m.bq_attr.push_back(attr - 1);
            for (int l = 0; l < Nfp; l++) {
                const int n1 = line_node(3, Np, r1, l, s1 ? N : 0);
                m.bq_n.push_back(n1);
                double nor[3];
                normal(e1, f1, n1, nor);
                for (int d = 0; d < 3; d++) m.bq_nor.push_back(nor[d]);
            }
        }
    }
    m.nf = (int)m.fe1.size();
    m.axis_idx.assign(NK, 0);
    m.attr_state.assign(5 * (size_t)std::max(q->nattr, 1), 0.0);
    for (int a = 0; a < q->nattr; a++)
        if (q->attr_state)
            for (int v = 0; v < 5; v++) m.attr_state[5 * (size_t)a + v] = q->attr_state[5 * (size_t)a + v];
    build_subcells(c);
    return 0;
}

/* DGSEMOperator::RecoverStateFromWeighted (DGSEMOperator.cpp:161-424) with
 * low_order_axis set, as an "axis" boundary sets it (Simulation.cpp:
 * 527-531), which makes every axis node "troubled" (line 238) and skips the
 * shape-function derivative.  Off the axis U = rU / r (lines 248-256).  On
 * it (in axis_idx, or r <= 0) each of rho, rho u_z, rho E is d(rU)/dr from
 * the two nearest off-axis nodes at the node's z (else the two nearest
 * overall): the least-squares slope through the origin, limited by minmod
 * against 1.5 f_j / r_j and capped at 10 |f_j1| / r_j1 (lines 302-378); one
 * such node gives its f / r; the radial momentum is 0.  A density below the
 * floor or a non-finite one or energy takes the nearest off-axis node's
 * state instead (lines 384-410). */
void axis_recover(const orc_ctx *c, const double *rU, double *U)
{
    const Mesh &m = c->m;
    const int K = m.K, NE = m.NE, nv = m.nv;
    const size_t NK = (size_t)NE * K;
    const double tiny = 1e-14, cap_mult = 10.0, z_tol = 1e-12, theta_lim = 1.5;
    const double rho_floor = c->p.axis_rho_floor > 0.0 ? c->p.axis_rho_floor : 1e-12;
    const double inf = std::numeric_limits<double>::infinity();
    auto sign = [](double a) { return (double)((a > 0) - (a < 0)); };
    auto minmod = [&](double a, double b) {
        if (a * b <= 0.0) return 0.0;
        return sign(a) * std::min(std::fabs(a), std::fabs(b));
    };
    auto clamp = [&](double x, double mx) {
        if (!std::isfinite(x)) return 0.0;
        return std::fabs(x) <= mx ? x : sign(x) * mx;
    };
    for (int e = 0; e < NE; e++) {
        const double *re = &m.X[NK + (size_t)e * K], *ze = &m.X[(size_t)e * K];
        for (int ld = 0; ld < K; ld++) {
            const size_t nd = (size_t)e * K + ld;
            const double r = re[ld];
            if (!m.axis_idx[nd] && r > 0.0) {
                for (int v = 0; v < nv; v++) U[v * NK + nd] = rU[v * NK + nd] / r;
                continue;
            }
            const double z_axis = ze[ld];
            auto d_dr = [&](int comp) -> double {
                auto f = [&](int j) { return rU[comp * NK + (size_t)e * K + j]; };
                int j1 = -1, j2 = -1;
                double r1 = inf, r2 = inf;
                for (int j = 0; j < K; j++) {
                    const double rj = re[j];
                    if (rj <= 0.0) continue;
                    if (std::fabs(ze[j] - z_axis) <= z_tol) {
                        if (rj < r1) { r2 = r1; j2 = j1; r1 = rj; j1 = j; }
                        else if (rj < r2) { r2 = rj; j2 = j; }
                    }
                }
                if (j1 < 0)
                    for (int j = 0; j < K; j++) {
                        const double rj = re[j];
                        if (rj > 0.0 && rj < r1) { r2 = r1; j2 = j1; r1 = rj; j1 = j; }
                        else if (rj > 0.0 && rj < r2) { r2 = rj; j2 = j; }
                    }
                const bool have_cap = j1 >= 0;
                const double f_ref = have_cap ? std::fabs(f(j1)) : 0.0;
                const double cap = (have_cap && !(f_ref < 1e-12)) ? cap_mult * f_ref / std::max(r1, tiny) : inf;
                if (j1 >= 0 && j2 >= 0 && std::isfinite(r1) && std::isfinite(r2) && r1 > 0.0 && r2 > 0.0) {
                    const double f1 = f(j1), f2 = f(j2);
                    const double mm = (r1 * f1 + r2 * f2) / std::max(r1 * r1 + r2 * r2, tiny);
                    const double s1 = f1 / r1, s2 = f2 / r2;
                    return clamp(minmod(mm, minmod(theta_lim * s1, theta_lim * s2)), cap);
                }
                if (j1 >= 0 && std::isfinite(r1) && r1 > 0.0) return clamp(f(j1) / r1, cap);
                return 0.0;
            };
            const double rho = d_dr(0), mz = d_dr(1), E = d_dr(3);
            if (!std::isfinite(rho) || rho < rho_floor || !std::isfinite(E)) {
                int nb = -1;
                double rn = -1.0;
                for (int j = 0; j < K; j++)
                    if (re[j] > 0.0 && (nb == -1 || re[j] < rn)) { rn = re[j]; nb = j; }
                /* create_impl refuses elements without an off-axis node */
                for (int v = 0; v < nv; v++) U[v * NK + nd] = rU[v * NK + (size_t)e * K + nb] / rn;
                U[2 * NK + nd] = 0.0;
            } else {
                U[0 * NK + nd] = rho;
                U[1 * NK + nd] = mz;
                U[2 * NK + nd] = 0.0;
                U[3 * NK + nd] = E;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* the RHS engine                                                       */
/* ------------------------------------------------------------------ */
/* n . x - offset - speed t for a split outer state fn = [kind, n0, n1, n2,
 * offset, speed, ...], summed left to right without contraction */
inline double split_side(const double *fn, const double *x, double t)
{
    double s = fn[1] * x[0];
    s = s + fn[2] * x[1];
    s = s + fn[3] * x[2];
    s = s - fn[4];
    s = s - fn[5] * t;
    return s;
}

template <class T> struct Engine {
    orc_ctx *c;
    const Mesh &m;
    int dim, N, Np, K, NE, nv, Nfp;
    size_t NK;
    double gamma;
    std::vector<T> U;        /* state copy (counted only) */
    /* entropy variables and gradients; every entry is written before it is
     * read, so they are allocated uninitialised and first touched by the
     * threads that fill them */
    std::unique_ptr<T[]> W, G;
    bool axi;
    std::vector<T> Urec;     /* axisymmetric: the state recovered from r U */

    explicit Engine(orc_ctx *cc) : c(cc), m(cc->m)
    {
        dim = m.dim; N = m.N; Np = m.Np; K = m.K; NE = m.NE; nv = m.nv; Nfp = m.Nfp;
        NK = (size_t)NE * K;
        gamma = c->p.gamma;
        axi = c->p.axisymmetric != 0;
    }

    /* r = y of the element map at reference point (xi, eta): the bilinear
     * map of the corners, as Tr.Transform (DGSEMIntegrator.cpp:272-276) */
    double rmap(int e, double xi, double eta) const
    {
        const double Nk[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta};
        double y = 0.0;
        for (int k = 0; k < 4; k++) y += Nk[k] * m.cq[(size_t)e * 8 + 2 * k + 1];
        return y;
    }
    /* r at the reference midpoint of nodes a and b (r_hat of lines 272-276
     * and 318-322, mid_r of ComputeFVFluxes lines 976-999) */
    double rmid(int e, int a, int b) const
    {
        const double *x = c->x.data();
        return rmap(e, 0.5 * (x[a % Np] + x[b % Np]), 0.5 * (x[a / Np] + x[b / Np]));
    }
    /* r at a face point: the node's coordinate (the face transformation's
     * point, lines 190-193 and BdrFaceIntegrator.cpp:81-83) */
    double rnode(int e, int n) const { return m.X[NK + (size_t)e * K + n]; }

    inline void state(const T *u, int e, int i, T *s) const
    {
        for (int v = 0; v < nv; v++) s[v] = u[v * NK + (size_t)e * K + i];
    }

    /* Indicator::CheckSmoothness (PerssonPeraireIndicator.cpp:21-58) and
     * DGSEMOperator::ComputeBlendingCoefficient (DGSEMOperator.cpp:60-80) */
    void blending(const T *u, int nth)
    {
        set_cat<T>(CAT_IND);
        const double s = 9.21024, thr = c->thresh;
        (void)nth;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
        {
        std::vector<T> q(K), md(K);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int e = 0; e < NE; e++) {
            for (int i = 0; i < K; i++) {
                T st[5];
                state(u, e, i, st);
                q[i] = st[0] * pressure(dim, gamma, st);
            }
            for (int a = 0; a < K; a++) {
                T y = c->Vinv[(size_t)a * K + 0] * q[0];
                for (int i = 1; i < K; i++) y += c->Vinv[(size_t)a * K + i] * q[i];
                md[a] = y;
            }
            T mm = 0.0, m1 = 0.0, m2 = 0.0;
            for (int a = 0; a < K; a++) {
                bool in1 = true, in2 = true;
                for (int d = 0; d < dim; d++) {
                    int dg = c->degs[a * dim + d];
                    if (dg > N - 2) { in2 = false; if (dg > N - 1) in1 = false; }
                }
                mm += md[a] * md[a];
                m1 += (in1 ? md[a] : T(0.0)) * (in1 ? md[a] : T(0.0));
                m2 += (in2 ? md[a] : T(0.0)) * (in2 ? md[a] : T(0.0));
            }
            T eta = 1.0 - m1 / mm;
            eta = mmax(eta, 1.0 - m2 / m1);
            c->eta[e] = val(eta);
            T al = 1.0 / (1.0 + mexp(-s * (eta - thr) / thr));
            double a = val(al);
            if (a < c->p.alpha_min) a = 0.0;
            else if (a > 1.0 - c->p.alpha_min) a = 1.0;
            a = std::min(a, c->p.alpha_max);
            c->alpha[e] = a;
        }
        }
    }

    /* DGSEMOperator.cpp:82-94, DGSEMNonlinearForm.cpp:424-645,
     * LiftingBR1.cpp:10-66 (faces) and 156-272 (elements),
     * DGSEMOperator.cpp:96-159 (gradient conversion) */
    void lifting(const T *u, int nth)
    {
        set_cat<T>(CAT_LIFT);
        W.reset(new T[nv * NK]);
        G.reset(new T[dim * nv * NK]);
        const double *D = c->D.data();
        const double w0 = c->w[0];
        (void)nth;
        /* Element loops are independent.  On boxes the interior faces of one
         * direction touch disjoint nodes and a node takes at most one face
         * term per direction, so running the directions in order, each in
         * parallel, adds every node's terms in the serial order.  The faces
         * of general meshes share corner nodes and run on one thread, and so
         * do the boundary faces. */
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
        {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int e = 0; e < NE; e++)
            for (int i = 0; i < K; i++) {
                T st[5], wv[5];
                state(u, e, i, st);
                cons2entropy(dim, gamma, st, wv);
                for (int v = 0; v < nv; v++) W[v * NK + (size_t)e * K + i] = wv[v];
            }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int e = 0; e < NE; e++)
            for (int id = 0; id < K; id++) {
                int ii[3] = {id % Np, (id / Np) % Np, id / (Np * Np)};
                T g1[5][3];
                for (int v = 0; v < nv; v++)
                    for (int r = 0; r < dim; r++) g1[v][r] = 0.0;
                for (int l = 0; l < Np; l++)
                    for (int r = 0; r < dim; r++) {
                        int jj[3] = {ii[0], ii[1], ii[2]};
                        jj[r] = l;
                        int nl = (jj[2] * Np + jj[1]) * Np + jj[0];
                        for (int v = 0; v < nv; v++)
                            g1[v][r] += W[v * NK + (size_t)e * K + nl] * D[ii[r] * Np + l];
                    }
                const double *adj = &m.adj[((size_t)e * K + id) * dim * dim];
                T invJ = 1.0 / T(m.J[(size_t)e * K + id]);
                for (int cc = 0; cc < dim; cc++)
                    for (int v = 0; v < nv; v++) {
                        T a = g1[v][0] * adj[0 * dim + cc];
                        for (int r = 1; r < dim; r++) a += g1[v][r] * adj[r * dim + cc];
                        G[(cc * nv + v) * NK + (size_t)e * K + id] = a * invJ;
                    }
            }
        auto face_terms = [&](int f) {
            for (int q = 0; q < Nfp; q++) {
                int e1 = m.fe1[f], e2 = m.fe2[f];
                if (e2 < 0) continue;
                int n1 = m.fn1[(size_t)f * Nfp + q], n2 = m.fn2[(size_t)f * Nfp + q];
                const double *nor = &m.fnor[((size_t)f * Nfp + q) * dim];
                T a1 = 1.0 / (w0 * T(m.J[(size_t)e1 * K + n1]));
                T a2 = 1.0 / (w0 * T(m.J[(size_t)e2 * K + n2]));
                for (int v = 0; v < nv; v++) {
                    T fv = 0.5 * (W[v * NK + (size_t)e2 * K + n2] - W[v * NK + (size_t)e1 * K + n1]);
                    for (int cc = 0; cc < dim; cc++) {
                        T fc = fv * nor[cc];
                        G[(cc * nv + v) * NK + (size_t)e1 * K + n1] += a1 * fc;
                        G[(cc * nv + v) * NK + (size_t)e2 * K + n2] += a2 * fc;
                    }
                }
            }
        };
        if (m.quad) {
#ifdef _OPENMP
#pragma omp single
#endif
            for (int f = 0; f < m.nf; f++) face_terms(f);
        } else {
            for (int r = 0; r < dim; r++) {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (int f = r * NE; f < (r + 1) * NE; f++) face_terms(f);
            }
        }
#ifdef _OPENMP
#pragma omp single
#endif
        {
        /* boundary faces, after the interior ones: LiftingBR1::
         * AssembleLiftingBdrFaceVector (LiftingBR1.cpp:311-407) adds
         * f n_c / (w_0 J) with the boundary type's lifting flux f */
        for (int r = 0; r < dim; r++)
            for (int side = 0; side < 2; side++) {
                const int type = c->p.bc[r][side];
                T outer[5];
                for (int v = 0; v < nv; v++) outer[v] = c->p.bc_state[r][side][v];
                for (int e : m.bel[r][side])
                    for (int q = 0; q < Nfp; q++) {
                        const int n1 = line_node(dim, Np, r, q, side ? N : 0);
                        T wv[5], f[5];
                        for (int v = 0; v < nv; v++) wv[v] = W[v * NK + (size_t)e * K + n1];
                        bc_lift_flux(dim, gamma, c->p.quirks, type, wv, outer, f);
                        T a1 = 1.0 / (w0 * T(m.J[(size_t)e * K + n1]));
                        const double nr = side ? m.ad[r] : -m.ad[r];
                        for (int v = 0; v < nv; v++)
                            for (int cc = 0; cc < dim; cc++) {
                                T fc = f[v] * (cc == r ? nr : 0.0);
                                G[(cc * nv + v) * NK + (size_t)e * K + n1] += a1 * fc;
                            }
                    }
            }
        /* the same on quadrilateral meshes, each point with its own normal */
        for (size_t b = 0; b < m.bq_e.size(); b++) {
            const int e = m.bq_e[b];
            T outer[5];
            for (int v = 0; v < nv; v++) outer[v] = m.attr_state[5 * (size_t)m.bq_attr[b] + v];
            for (int q = 0; q < Nfp; q++) {
                const int n1 = m.bq_n[b * Nfp + q];
                T wv[5], f[5];
                for (int v = 0; v < nv; v++) wv[v] = W[v * NK + (size_t)e * K + n1];
                bc_lift_flux(dim, gamma, c->p.quirks, m.bq_type[b], wv, outer, f);
                T a1 = 1.0 / (w0 * T(m.J[(size_t)e * K + n1]));
                for (int v = 0; v < nv; v++)
                    for (int cc = 0; cc < dim; cc++) {
                        T fc = f[v] * m.bq_nor[(b * Nfp + q) * dim + cc];
                        G[(cc * nv + v) * NK + (size_t)e * K + n1] += a1 * fc;
                    }
            }
        }
        }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int e = 0; e < NE; e++)
            for (int i = 0; i < K; i++) {
                T st[5];
                state(u, e, i, st);
                for (int cc = 0; cc < dim; cc++) {
                    T g[5];
                    for (int v = 0; v < nv; v++) g[v] = G[(cc * nv + v) * NK + (size_t)e * K + i];
                    entgrad2primgrad(dim, gamma, c->p.quirks && !(c->p.plant & ORC_PLANT_GRAD_NO_ALIAS), st, g);
                    for (int v = 0; v < nv; v++) G[(cc * nv + v) * NK + (size_t)e * K + i] = g[v];
                }
            }
        }
    }

    /* DGSEMIntegrator::ComputeFVFluxes, DGSEMIntegrator.cpp:972-1196.
     * fv (K x nv, node-major) is overwritten with alpha * FV. */
    T fv_element(const T *u, int e, T alpha_e, T *fv, T lam)
    {
        set_cat<T>(CAT_FV);
        const int q = c->p.quirks;
        for (int i = 0; i < K * nv; i++) fv[i] = 0.0;
        for (int r = 0; r < dim; r++)
            for (int l = 0; l < Nfp; l++) {
                T dU[5], s1[5], s2[5], F[5];
                for (int v = 0; v < nv; v++) dU[v] = 0.0;
                int id1 = line_node(dim, Np, r, l, 0);
                state(u, e, id1, s1);
                for (int i = 0; i < Np - 1; i++) {
                    int id2 = line_node(dim, Np, r, l, i + 1);
                    state(u, e, id2, s2);
                    T nor[3];
                    const double *sn = &m.sub[((((size_t)e * dim + r) * Nfp + l) * N + i) * dim];
                    for (int d = 0; d < dim; d++) nor[d] = sn[d];
                    lam = mmax(lam, face_flux(dim, gamma, q, s1, s2, nor, F));
                    if (axi) {
                        /* flux_num *= mid_r (lines 1030-1032, 1110-1112) */
                        const T rh = rmid(e, id1, id2);
                        for (int v = 0; v < nv; v++) F[v] *= rh;
                    }
                    T sc = 1.0 / (T(m.J[(size_t)e * K + id1]) * c->w[i]);
                    for (int v = 0; v < nv; v++) {
                        dU[v] -= F[v];
                        dU[v] *= sc;
                        if (r == 0) fv[id1 * nv + v] = dU[v];
                        else fv[id1 * nv + v] += dU[v];
                        dU[v] = F[v];
                        s1[v] = s2[v];
                    }
                    id1 = id2;
                }
                T sc = 1.0 / (T(m.J[(size_t)e * K + id1]) * c->w[Np - 1]);
                for (int v = 0; v < nv; v++) {
                    dU[v] *= sc;
                    if (r == 0) fv[id1 * nv + v] = dU[v];
                    else fv[id1 * nv + v] += dU[v];
                }
            }
        for (int i = 0; i < K * nv; i++) fv[i] *= alpha_e;
        return lam;
    }

    /* AssembleElementVector: Euler DGSEMIntegrator.cpp:224-433,
     * NS-3D 601-750, NS-2D 752-864.  Adds the element vector to dudt. */
    T element(const T *u, int e, T *dudt, T lam, std::vector<T> &Fi, std::vector<T> &Fv,
              std::vector<T> &el)
    {
        const bool par = c->p.parabolic != 0, blend = c->p.blending != 0;
        const int quirks = c->p.quirks;
        const T alpha_e = blend ? T(c->alpha[e]) : T(0.0);
        if (blend) lam = fv_element(u, e, alpha_e, el.data(), lam);
        else for (int i = 0; i < K * nv; i++) el[i] = 0.0;
        /* viscous volume flux, contravariant (DGSEMIntegrator.cpp:617-641) */
        if (par) {
            set_cat<T>(CAT_VISC);
            for (int i = 0; i < K; i++) {
                /* stale member id1 (line 620): the last node of the previous
                 * element loop, or of ComputeFVFluxes' zeta sweep (lines
                 * 1165-1190), both K-1 */
                const int stale = (c->p.plant & ORC_PLANT_STALE_NODE0) ? 0 : K - 1;
                int src = quirks ? stale : i;
                T st[5], dq[3][5], fl[5 * 3];
                state(u, e, src, st);
                const T *dqp[3];
                for (int cc = 0; cc < dim; cc++) {
                    for (int v = 0; v < nv; v++) dq[cc][v] = G[(cc * nv + v) * NK + (size_t)e * K + i];
                    dqp[cc] = dq[cc];
                }
                viscous_flux(dim, gamma, c->p.Pr, c->p.mu, c->p.sutherland ? c->p.R_gas : 0.0, st, dqp, fl);
                const double *adj = &m.adj[((size_t)e * K + i) * dim * dim];
                for (int r = 0; r < dim; r++)
                    for (int v = 0; v < nv; v++) {
                        T a = adj[r * dim + 0] * fl[v * dim + 0];
                        for (int cc = 1; cc < dim; cc++) a += adj[r * dim + cc] * fl[v * dim + cc];
                        Fv[((size_t)r * K + i) * nv + v] = a;
                    }
            }
        }
        /* inviscid pairs, node loop (k, j, i) as the reference */
        const int Nz = dim == 3 ? Np : 1, Ny = dim >= 2 ? Np : 1;
        auto FI = [&](int r, int node, int part) -> T * {
            return &Fi[(((size_t)r * K + node) * Np + part) * nv];
        };
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Np; i++) {
                    set_cat<T>(CAT_VOL);
                    int id1 = (k * Np + j) * Np + i;
                    int ii[3] = {i, j, k};
                    T s1[5], s2[5], m1[3], m2[3];
                    state(u, e, id1, s1);
                    const double *adj1 = &m.adj[((size_t)e * K + id1) * dim * dim];
                    T glast[5];
                    for (int r = 0; r < dim; r++) {
                        for (int d = 0; d < dim; d++) m1[d] = adj1[r * dim + d];
                        T f[5];
                        for (int v = 0; v < nv; v++) f[v] = 0.0;
                        T *self = FI(r, id1, ii[r]);
                        for (int v = 0; v < nv; v++) self[v] = 0.0;
                        for (int mm = ii[r] + 1; mm < Np; mm++) {
                            int jj[3] = {i, j, k};
                            jj[r] = mm;
                            int id2 = (jj[2] * Np + jj[1]) * Np + jj[0];
                            state(u, e, id2, s2);
                            const double *adj2 = &m.adj[((size_t)e * K + id2) * dim * dim];
                            for (int d = 0; d < dim; d++) m2[d] = adj2[r * dim + d];
                            lam = mmax(lam, volume_flux(dim, gamma, s1, s2, m1, m2, f));
                            if (axi) {
                                /* f *= r_hat (lines 290-292, 337-339) */
                                const T rh = rmid(e, id1, id2);
                                for (int v = 0; v < nv; v++) f[v] *= rh;
                            }
                            T *a = FI(r, id1, mm), *b = FI(r, id2, ii[r]);
                            for (int v = 0; v < nv; v++) {
                                a[v] = f[v];
                                /* NS-3D: H_inviscid(id2).SetCol(k, g) (line 702) */
                                b[v] = (par && quirks && dim == 3 && r == 2 && !(c->p.plant & ORC_PLANT_ZETA_F))
                                           ? glast[v] : f[v];
                            }
                        }
                        if (r == ((c->p.plant & ORC_PLANT_ZETA_XI) ? 0 : 1))
                            for (int v = 0; v < nv; v++) glast[v] = f[v];
                    }
                    set_cat<T>(CAT_ASM);
                    T dU[5], dV[5];
                    for (int v = 0; v < nv; v++) { dU[v] = 0.0; dV[v] = 0.0; }
                    for (int r = 0; r < dim; r++) {
                        const double *Drow = &c->Dh2[ii[r] * Np];
                        for (int mm = 0; mm < Np; mm++) {
                            const T *fp = FI(r, id1, mm);
                            for (int v = 0; v < nv; v++) {
                                if (r == 0 && mm == 0) dU[v] = Drow[mm] * fp[v];
                                else dU[v] += Drow[mm] * fp[v];
                            }
                        }
                        if (par) {
                            const double *Dr = &c->Dh[ii[r] * Np];
                            for (int l = 0; l < Np; l++) {
                                int jj[3] = {i, j, k};
                                jj[r] = l;
                                int nl = (jj[2] * Np + jj[1]) * Np + jj[0];
                                const T *fp = &Fv[((size_t)r * K + nl) * nv];
                                for (int v = 0; v < nv; v++) dV[v] += fp[v] * Dr[l];
                            }
                        }
                    }
                    T invJ = 1.0 / T(m.J[(size_t)e * K + id1]);
                    if (axi) {
                        /* the radial momentum's pressure source, before the
                         * volume term (lines 411-426, PressureFromConservative
                         * lines 1306-1315) */
                        const T ke = 0.5 * ((s1[1] * s1[1] + s1[2] * s1[2]) / s1[0]);
                        el[id1 * nv + 2] += (gamma - 1.0) * (s1[3] - ke);
                    }
                    for (int v = 0; v < nv; v++) {
                        T a = -dU[v];
                        if (blend) a *= (1.0 - alpha_e);
                        if (par) a = a + dV[v];
                        a *= invJ;
                        el[id1 * nv + v] += a;
                    }
                }
        for (int i = 0; i < K; i++)
            for (int v = 0; v < nv; v++) dudt[v * NK + (size_t)e * K + i] += el[i * nv + v];
        return lam;
    }

    /* interior faces of direction r, faces [f0, f1):
     * AssembleFaceVector Euler DGSEMIntegrator.cpp:151-222,
     * NS-3D 435-494, NS-2D 496-549 */
    T faces(const T *u, int f0, int f1, T *dudt, T lam)
    {
        const bool par = c->p.parabolic != 0;
        const int quirks = c->p.quirks, ec = c->p.ec_surface;
        const double w0 = c->w[0];
        for (int f = f0; f < f1; f++)
            for (int q = 0; q < Nfp; q++) {
                int e1 = m.fe1[f], e2 = m.fe2[f];
                if (e2 < 0) continue;
                int n1 = m.fn1[(size_t)f * Nfp + q], n2 = m.fn2[(size_t)f * Nfp + q];
                T s1[5], s2[5], nor[3], F[5];
                state(u, e1, n1, s1);
                state(u, e2, n2, s2);
                for (int d = 0; d < dim; d++) nor[d] = m.fnor[((size_t)f * Nfp + q) * dim + d];
                T d1[5], d2[5], fn[5];
                if (par) {
                    /* viscous fluxes first, as DGSEMIntegrator.cpp:467-477 */
                    set_cat<T>(CAT_VISC);
                    T fl1[15], fl2[15], dq[3][5];
                    const T *dqp[3];
                    for (int cc = 0; cc < dim; cc++) {
                        for (int v = 0; v < nv; v++) dq[cc][v] = G[(cc * nv + v) * NK + (size_t)e1 * K + n1];
                        dqp[cc] = dq[cc];
                    }
                    viscous_flux(dim, gamma, c->p.Pr, c->p.mu, c->p.sutherland ? c->p.R_gas : 0.0, s1, dqp, fl1);
                    for (int cc = 0; cc < dim; cc++)
                        for (int v = 0; v < nv; v++) dq[cc][v] = G[(cc * nv + v) * NK + (size_t)e2 * K + n2];
                    viscous_flux(dim, gamma, c->p.Pr, c->p.mu, c->p.sutherland ? c->p.R_gas : 0.0, s2, dqp, fl2);
                    for (int v = 0; v < nv; v++) {
                        fn[v] = 0.0;
                        for (int cc = 0; cc < dim; cc++) {
                            T avg = (fl1[v * dim + cc] + fl2[v * dim + cc]) * 0.5;
                            fn[v] = (cc == 0) ? avg * nor[0] : fn[v] + avg * nor[cc];
                        }
                    }
                    if (quirks && !(c->p.plant & ORC_PLANT_FACE_NO_ALIAS)) {
                        /* ComputeViscousFlux -> Conserv2Prim -> ComputePressure
                         * (NavierStokesFlux.cpp:14, BasicOperations.cpp:33-34,185)
                         * divides the momentum slots of state1/state2 in place,
                         * and ComputeFaceFlux (line 481) then reads them */
                        T r1 = 1.0 / s1[0], r2 = 1.0 / s2[0];
                        for (int d = 0; d < dim; d++) { s1[1 + d] *= r1; s2[1 + d] *= r2; }
                    }
                }
                set_cat<T>(CAT_FACE);
                if (ec) lam = mmax(lam, volume_flux(dim, gamma, s1, s2, nor, nor, F));
                else lam = mmax(lam, face_flux(dim, gamma, quirks, s1, s2, nor, F));
                if (axi) {
                    /* flux_num *= r (lines 190-205) */
                    const T rf = rnode(e1, n1);
                    for (int v = 0; v < nv; v++) F[v] *= rf;
                }
                for (int v = 0; v < nv; v++) { d1[v] = -F[v]; d2[v] = F[v]; }
                if (par)
                    for (int v = 0; v < nv; v++) { d1[v] += fn[v]; d2[v] -= fn[v]; }
                set_cat<T>(CAT_ASM);
                T a1 = 1.0 / (w0 * T(m.J[(size_t)e1 * K + n1]));
                T a2 = 1.0 / (w0 * T(m.J[(size_t)e2 * K + n2]));
                for (int v = 0; v < nv; v++) {
                    dudt[v * NK + (size_t)e1 * K + n1] += a1 * d1[v];
                    dudt[v * NK + (size_t)e2 * K + n2] += a2 * d2[v];
                }
            }
        return lam;
    }

    /* boundary faces of side `side` of direction r, elements [b0, b1) of
     * that side's list: BdrFaceIntegrator::AssembleFaceVector
     * (BdrFaceIntegrator.cpp:61-123) adds -F_b / (w_0 J) at the face node of
     * the interior element, F_b the boundary type's flux for the outward
     * normal (CalcOrtho of the face Jacobian: -/+ |face| e_r) */
    T bfaces(const T *u, int r, int side, int b0, int b1, T *dudt, T lam)
    {
        const int type = c->p.bc[r][side];
        const double w0 = c->w[0];
        T outer[5], nor[3];
        for (int v = 0; v < nv; v++) outer[v] = c->p.bc_state[r][side][v];
        for (int d = 0; d < dim; d++) nor[d] = 0.0;
        nor[r] = side ? m.ad[r] : -m.ad[r];
        for (int b = b0; b < b1; b++) {
            const int e = m.bel[r][side][b];
            for (int q = 0; q < Nfp; q++) {
                const int n1 = line_node(dim, Np, r, q, side ? N : 0);
                T s1[5], F[5];
                state(u, e, n1, s1);
                set_cat<T>(CAT_FACE);
                lam = mmax(lam, bc_flux(dim, gamma, c->p.quirks, type, s1, outer, nor, F));
                T d[5];
                for (int v = 0; v < nv; v++) d[v] = -F[v];
                if (c->p.parabolic) {
                    /* ComputeBdrFaceViscousFlux (BdrFaceIntegrator.cpp:141-175):
                     * -F* + Fv.n; Fv of the interior state and gradients
                     * (outflow, wall) or of the outer state with zero
                     * gradients (inflow, specified state); the adiabatic wall
                     * replaces the energy term by V.(Fv.n)_mom + qn |n| */
                    set_cat<T>(CAT_VISC);
                    T fl[15], dq[3][5], zero[5], fn[5];
                    const T *dqp[3];
                    const bool ext = type == ORC_BC_SUPERSONIC_INFLOW || type == ORC_BC_SPECIFIED_STATE;
                    for (int v = 0; v < nv; v++) zero[v] = 0.0;
                    for (int cc = 0; cc < dim; cc++) {
                        for (int v = 0; v < nv; v++) dq[cc][v] = G[(cc * nv + v) * NK + (size_t)e * K + n1];
                        dqp[cc] = ext ? zero : dq[cc];
                    }
                    T sv[5];
                    for (int v = 0; v < nv; v++) sv[v] = ext ? outer[v] : s1[v];
                    viscous_flux(dim, gamma, c->p.Pr, c->p.mu, c->p.sutherland ? c->p.R_gas : 0.0, sv, dqp, fl);
                    for (int v = 0; v < nv; v++) {
                        fn[v] = 0.0;
                        for (int cc = 0; cc < dim; cc++) fn[v] += fl[v * dim + cc] * nor[cc];
                    }
                    if (type == ORC_BC_NO_SLIP_ADIABATIC) {
                        T nn = 0.0;
                        for (int cc = 0; cc < dim; cc++) nn += nor[cc] * nor[cc];
                        T qn = T(c->p.bc_state[r][side][3]) * msqrt(nn);
                        T en = T(c->p.bc_state[r][side][0]) * fn[1];
                        for (int cc = 1; cc < dim; cc++) en = en + T(c->p.bc_state[r][side][cc]) * fn[1 + cc];
                        fn[nv - 1] = en + qn;
                    }
                    for (int v = 0; v < nv; v++) d[v] += fn[v];
                }
                set_cat<T>(CAT_ASM);
                T a1 = 1.0 / (w0 * T(m.J[(size_t)e * K + n1]));
                for (int v = 0; v < nv; v++) dudt[v * NK + (size_t)e * K + n1] += a1 * d[v];
            }
        }
        return lam;
    }

    /* boundary faces of a quadrilateral mesh, [b0, b1) of its list: as
     * bfaces, with each face's own normal and attribute */
    T bfaces_quad(const T *u, int b0, int b1, T *dudt, T lam)
    {
        const double w0 = c->w[0];
        for (int b = b0; b < b1; b++) {
            const int e = m.bq_e[b];
            T outer[5];
            for (int v = 0; v < nv; v++) outer[v] = m.attr_state[5 * (size_t)m.bq_attr[b] + v];
            const double *fn = (c->attr_fn.size() > 16 * (size_t)m.bq_attr[b] &&
                                c->attr_fn[16 * (size_t)m.bq_attr[b]] != 0.0)
                                   ? &c->attr_fn[16 * (size_t)m.bq_attr[b]] : nullptr;
            for (int q = 0; q < Nfp; q++) {
                const int n1 = m.bq_n[(size_t)b * Nfp + q];
                if (fn) {
                    /* the time-dependent outer state at this boundary point
                     * (conserv_state_fun.SetTime(time); Eval at the point,
                     * SpecifiedStateBdrFaceIntegrator.cpp:24-34): fn's "below"
                     * state where n . x - offset - speed t < 0 */
                    double x[3] = {0.0, 0.0, 0.0};
                    for (int d = 0; d < dim; d++) x[d] = m.X[d * NK + (size_t)e * K + n1];
                    const double sgn = split_side(fn, x, c->t_stage);
                    for (int v = 0; v < nv; v++) outer[v] = sgn < 0.0 ? fn[6 + v] : fn[11 + v];
                }
                T s1[5], F[5], nor[3];
                state(u, e, n1, s1);
                for (int d = 0; d < dim; d++) nor[d] = m.bq_nor[((size_t)b * Nfp + q) * dim + d];
                set_cat<T>(CAT_FACE);
                const int type = m.bq_type[b];
                lam = mmax(lam, bc_flux(dim, gamma, c->p.quirks, type, s1, outer, nor, F));
                if (axi) {
                    /* dU_face1 *= r (BdrFaceIntegrator.cpp:109-111) */
                    const T rf = rnode(e, n1);
                    for (int v = 0; v < nv; v++) F[v] *= rf;
                }
                T d[5];
                for (int v = 0; v < nv; v++) d[v] = -F[v];
                if (c->p.parabolic) {
                    /* as bfaces: -F* + Fv.n */
                    set_cat<T>(CAT_VISC);
                    T fl[15], dq[3][5], zero[5], fn[5], sv[5];
                    const T *dqp[3];
                    const bool ext = type == ORC_BC_SUPERSONIC_INFLOW || type == ORC_BC_SPECIFIED_STATE;
                    for (int v = 0; v < nv; v++) { zero[v] = 0.0; sv[v] = ext ? outer[v] : s1[v]; }
                    for (int cc = 0; cc < dim; cc++) {
                        for (int v = 0; v < nv; v++) dq[cc][v] = G[(cc * nv + v) * NK + (size_t)e * K + n1];
                        dqp[cc] = ext ? zero : dq[cc];
                    }
                    viscous_flux(dim, gamma, c->p.Pr, c->p.mu, c->p.sutherland ? c->p.R_gas : 0.0, sv, dqp, fl);
                    for (int v = 0; v < nv; v++) {
                        fn[v] = 0.0;
                        for (int cc = 0; cc < dim; cc++) fn[v] += fl[v * dim + cc] * nor[cc];
                    }
                    if (type == ORC_BC_NO_SLIP_ADIABATIC) {
                        T nn = 0.0;
                        for (int cc = 0; cc < dim; cc++) nn += nor[cc] * nor[cc];
                        T qn = outer[3] * msqrt(nn);
                        T en = outer[0] * fn[1];
                        for (int cc = 1; cc < dim; cc++) en = en + outer[cc] * fn[1 + cc];
                        fn[nv - 1] = en + qn;
                    }
                    for (int v = 0; v < nv; v++) d[v] += fn[v];
                }
                set_cat<T>(CAT_ASM);
                T a1 = 1.0 / (w0 * T(m.J[(size_t)e * K + n1]));
                for (int v = 0; v < nv; v++) dudt[v * NK + (size_t)e * K + n1] += a1 * d[v];
            }
        }
        return lam;
    }

    /* DGSEMOperator::Mult (DGSEMOperator.cpp:537-591) with
     * DGSEMNonlinearForm::Mult (DGSEMNonlinearForm.cpp:647-849, 1302-1545):
     * dudt is overwritten. */
    void rhs(const T *u, T *dudt, int threads)
    {
        if (axi) {
            /* RecoverStateFromWeighted(u, U) first (DGSEMOperator.cpp:539-541) */
            std::vector<double> ud(nv * NK), Ud(nv * NK);
            for (size_t i = 0; i < nv * NK; i++) ud[i] = val(u[i]);
            axis_recover(c, ud.data(), Ud.data());
            Urec.resize(nv * NK);
            for (size_t i = 0; i < nv * NK; i++) Urec[i] = T(Ud[i]);
            u = Urec.data();
        }
        int nth = threads > 0 ? threads : 1;
        if (c->p.blending) blending(u, nth);
        if (c->p.parabolic) lifting(u, nth);
        for (size_t i = 0; i < nv * NK; i++) dudt[i] = 0.0;
        double lam = c->lam;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth) reduction(max : lam)
#endif
        {
            std::vector<T> Fi((size_t)dim * K * Np * nv), Fv((size_t)dim * K * nv), el((size_t)K * nv);
            T l = lam;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int e = 0; e < NE; e++) l = element(u, e, dudt, l, Fi, Fv, el);
            if (m.quad) {
                /* faces of a general mesh share corner nodes: one thread */
#ifdef _OPENMP
#pragma omp single
#endif
                {
                    l = faces(u, 0, m.nf, dudt, l);
                    l = bfaces_quad(u, 0, (int)m.bq_e.size(), dudt, l);
                }
            }
            /* faces of one direction touch disjoint nodes of each element */
            for (int r = 0; r < dim && !m.quad; r++) {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (int b = 0; b < NE; b++) l = faces(u, r * NE + b, r * NE + b + 1, dudt, l);
            }
            /* boundary faces after the interior ones (DGSEMNonlinearForm.cpp:776-826);
             * the faces of one side touch disjoint nodes */
            for (int r = 0; r < dim && !m.quad; r++)
                for (int side = 0; side < 2; side++) {
                    const int nb = (int)m.bel[r][side].size();
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                    for (int b = 0; b < nb; b++) l = bfaces(u, r, side, b, b + 1, dudt, l);
                }
            lam = std::max(lam, val(l));
        }
        (void)nth;
        c->lam = lam;
        if (axi)
            /* ZeroAxisRadialMom(dudt) (DGSEMOperator.cpp:573-583, 521-533) */
            for (size_t i = 0; i < NK; i++)
                if (m.axis_idx[i]) dudt[2 * NK + i] = 0.0;
    }
};

/* MFEM RK3SSPSolver::Step (Shu-Osher form), called at Simulation.cpp:1010 */
void rk3_step(orc_ctx *c, double *x, double dt)
{
    Engine<double> E(c);
    const size_t n = (size_t)c->m.nv * c->m.NE * c->m.K;
    std::vector<double> &y = c->stage_y, &k = c->stage_k;
    y.resize(n); k.resize(n);
    /* stage times t, t + dt, t + dt/2 (RK3SSPSolver's SetTime calls) */
    const double t = c->time;
    c->t_stage = t;
    E.rhs(x, k.data(), c->p.threads);
    for (size_t i = 0; i < n; i++) y[i] = x[i] + dt * k[i];
    c->t_stage = t + dt;
    E.rhs(y.data(), k.data(), c->p.threads);
    for (size_t i = 0; i < n; i++) y[i] += dt * k[i];
    for (size_t i = 0; i < n; i++) y[i] = 3. / 4 * x[i] + 1. / 4 * y[i];
    c->t_stage = t + dt / 2;
    E.rhs(y.data(), k.data(), c->p.threads);
    for (size_t i = 0; i < n; i++) y[i] += dt * k[i];
    for (size_t i = 0; i < n; i++) x[i] = 1. / 3 * x[i] + 2. / 3 * y[i];
    c->time = t + dt;
}

} /* namespace */

/* ------------------------------------------------------------------ */
/* C ABI                                                                */
/* ------------------------------------------------------------------ */
extern "C" {

int orc_lgl(int order, double *x, double *w)
{
    if (order < 1 || order > 16) return 1;
    std::vector<double> xv, wv;
    lgl_rule(order, xv, wv);
    std::memcpy(x, xv.data(), xv.size() * sizeof(double));
    std::memcpy(w, wv.data(), wv.size() * sizeof(double));
    return 0;
}

int orc_dmat(int order, double *D, double *Dhat, double *Dhat2)
{
    if (order < 1 || order > 16) return 1;
    std::vector<double> x, w, a, b, cc;
    lgl_rule(order, x, w);
    dmats(order, x, w, a, b, cc);
    std::memcpy(D, a.data(), a.size() * sizeof(double));
    std::memcpy(Dhat, b.data(), b.size() * sizeof(double));
    std::memcpy(Dhat2, cc.data(), cc.size() * sizeof(double));
    return 0;
}

int orc_legendre_vdm(int order, double *V1, double *V1inv)
{
    if (order < 1 || order > 16) return 1;
    const int Np = order + 1;
    std::vector<double> x, w, V(Np * Np), Vi;
    lgl_rule(order, x, w);
    std::vector<double> L(Np);
    for (int i = 0; i < Np; i++) {
        calc_legendre(order, x[i], L.data());
        for (int a = 0; a < Np; a++) V[i * Np + a] = L[a];
    }
    if (!invert(Np, V, Vi)) return 2;
    std::memcpy(V1, V.data(), V.size() * sizeof(double));
    std::memcpy(V1inv, Vi.data(), Vi.size() * sizeof(double));
    return 0;
}

double orc_logmean(double x, double y) { return logmean<double>(x, y); }

double orc_volume_flux(int dim, double gamma, const double *s1, const double *s2,
                       const double *m1, const double *m2, double *F)
{
    return volume_flux<double>(dim, gamma, s1, s2, m1, m2, F);
}

double orc_face_flux(int dim, double gamma, int quirks, const double *s1, const double *s2,
                     const double *nor, double *F)
{
    return face_flux<double>(dim, gamma, quirks, s1, s2, nor, F);
}

void orc_cons2entropy(int dim, double gamma, const double *s, double *w)
{
    cons2entropy<double>(dim, gamma, s, w);
}

void orc_entgrad2primgrad(int dim, double gamma, int quirks, const double *s, double *g)
{
    entgrad2primgrad<double>(dim, gamma, quirks, s, g);
}

void orc_viscous_flux(int dim, double gamma, double Pr, double mu, double R_gas, const double *s,
                      const double *dqdx, const double *dqdy, const double *dqdz, double *flux)
{
    const double *dq[3] = {dqdx, dqdy, dqdz};
    viscous_flux<double>(dim, gamma, Pr, mu, R_gas, s, dq, flux);
}

double orc_bc_flux(int dim, double gamma, int quirks, int type, const double *s1, const double *outer,
                   const double *nor, double *F)
{
    return bc_flux<double>(dim, gamma, quirks, type, s1, outer, nor, F);
}

void orc_bc_lift_flux(int dim, double gamma, int quirks, int type, const double *w, const double *outer,
                      double *f)
{
    bc_lift_flux<double>(dim, gamma, quirks, type, w, outer, f);
}

static int create_impl(const orc_params *p, const orc_quad_mesh *qm, orc_ctx **out,
                       const orc_hex_mesh *hm = nullptr);

int orc_create(const orc_params *p, orc_ctx **out) { return create_impl(p, nullptr, out); }

int orc_create_hex(const orc_params *p, const orc_hex_mesh *m, orc_ctx **out)
{
    if (!p || !m || p->dim != 3) return 1;
    if (m->nelem < 1 || !m->elem || (!m->vert && !m->elem_coords)) return 3;
    return create_impl(p, nullptr, out, m);
}

int orc_create_quad(const orc_params *p, const orc_quad_mesh *m, orc_ctx **out)
{
    if (!p || !m || !out || p->dim != 2) return 1;
    for (int a = 0; a < m->nattr; a++) {
        const int t = m->attr_bc[a];
        if (p->parabolic && (t == ORC_BC_SLIP || t == ORC_BC_SYMMETRY)) return 7;
        if (!p->parabolic && t == ORC_BC_NO_SLIP_ADIABATIC) return 7;
        if (p->quirks && t == ORC_BC_NO_SLIP_ADIABATIC && m->attr_state && m->attr_state[4 * (size_t)a + 3] != 0.0)
            return 7;
    }
    return create_impl(p, m, out);
}

static int create_impl(const orc_params *p, const orc_quad_mesh *qm, orc_ctx **out, const orc_hex_mesh *hm)
{
    if (!p || !out) return 1;
    if (p->dim < 1 || p->dim > 3 || p->order < 1 || p->order > 12) return 2;
    /* 1-D: the Euler equations on a segment (every 1-D case of the
     * reference, TestCases/Euler/1D, is an Euler case) */
    if (p->dim == 1 && (p->parabolic || qm || hm)) return 2;
    const bool mesh = qm || hm;
    for (int r = 0; r < p->dim && !mesh; r++)
        if (p->nelem[r] < 1 || !(p->hi[r] > p->lo[r])) return 3;
    if (p->sutherland && !(p->R_gas > 0.0)) return 5;
    /* axisymmetric: 2-D quadrilateral meshes, Euler (the reference's
     * Navier-Stokes element and face routines carry no r weighting) */
    if (p->axisymmetric && (!qm || p->parabolic)) return 9;
    for (int r = 0; r < p->dim && !mesh; r++) {
        for (int side = 0; side < 2; side++)
            if (p->bc[r][side] < ORC_BC_PERIODIC || p->bc[r][side] > ORC_BC_NO_SLIP_ADIABATIC) return 6;
        if ((p->bc[r][0] == ORC_BC_PERIODIC) != (p->bc[r][1] == ORC_BC_PERIODIC)) return 6;
        for (int side = 0; side < 2; side++) {
            const int t = p->bc[r][side];
            /* the reference has no viscous or lifting flux for slip and
             * symmetry sides (BdrFaceIntegrator.cpp:247-270 raise), and
             * no-slip walls without viscous terms are not used */
            if (p->parabolic && (t == ORC_BC_SLIP || t == ORC_BC_SYMMETRY)) return 7;
            if (!p->parabolic && t == ORC_BC_NO_SLIP_ADIABATIC) return 7;
            /* NoSlipAdiabWallBdrFaceIntegrator.cpp:27 scales its member qn by
             * |n| on every call: only qn = 0 is reproducible */
            if (p->quirks && t == ORC_BC_NO_SLIP_ADIABATIC && p->bc_state[r][side][3] != 0.0) return 7;
        }
    }
    orc_ctx *c = new (std::nothrow) orc_ctx();
    if (!c) return 4;
    c->p = *p;
    lgl_rule(p->order, c->x, c->w);
    dmats(p->order, c->x, c->w, c->D, c->Dh, c->Dh2);
    if (hm) {
        if (build_mesh_hex(c, hm) != 0) {
            delete c;
            return 8;
        }
    } else if (qm) {
        if (build_mesh_quad(c, qm) != 0) {
            delete c;
            return 8;
        }
    } else {
        build_mesh(c);
    }
    const Mesh &m = c->m;
    if (p->axisymmetric) {
        /* nodes on the axis need an axis boundary (low_order_axis; the
         * shape-derivative path depends on the previous call's alpha), none
         * may lie below it, and every element keeps an off-axis node */
        const size_t NK = (size_t)m.NE * m.K;
        for (int e = 0; e < m.NE; e++) {
            bool off = false, on = false;
            for (int i = 0; i < m.K; i++) {
                const double r = m.X[NK + (size_t)e * m.K + i];
                if (r < -1e-14) { delete c; return 9; }
                if (r > 0.0) off = true;
                if (!(r > 0.0) || m.axis_idx[(size_t)e * m.K + i]) on = true;
            }
            if (!off || (on && !m.axis_bc)) { delete c; return 9; }
        }
    }
    c->alpha.assign(m.NE, 0.0);
    c->eta.assign(m.NE, 0.0);
    c->lam = -1.0; /* DGSEMIntegrator.cpp:39 */
    const double tdeg = p->order + ((p->plant & ORC_PLANT_THRESH_NP1) ? 1 : 0);
    c->thresh = 0.5 * std::pow(10.0, -1.8 * std::pow(tdeg, 0.25)); /* DGSEMOperator.cpp:25 */
    /* ModalBasis (ModalBasis.cpp:42-135): mode n <-> per-dim degrees,
     * V(node, n) = prod_d P_deg(x_d); V^-1 = kron of the 1-D inverses */
    const int Np = m.Np, K = m.K, dim = m.dim;
    c->degs.assign(K * dim, 0);
    for (int n = 0; n < K; n++) {
        if (dim == 1) c->degs[n] = n;
        else if (dim == 2) { c->degs[n * 2 + 0] = n / Np; c->degs[n * 2 + 1] = n % Np; }
        else {
            c->degs[n * 3 + 0] = n / (Np * Np);
            c->degs[n * 3 + 1] = (n / Np) % Np;
            c->degs[n * 3 + 2] = n % Np;
        }
    }
    std::vector<double> V1(Np * Np), V1i(Np * Np);
    orc_legendre_vdm(p->order, V1.data(), V1i.data());
    c->Vinv.assign((size_t)K * K, 0.0);
    for (int n = 0; n < K; n++)
        for (int i = 0; i < K; i++) {
            int id[3] = {i % Np, (i / Np) % Np, i / (Np * Np)};
            double v = 1.0;
            for (int d = 0; d < dim; d++) v *= V1i[c->degs[n * dim + d] * Np + id[d]];
            c->Vinv[(size_t)n * K + i] = v;
        }
    c->wJ.assign((size_t)m.NE * K, 0.0);
    for (int e = 0; e < m.NE; e++)
        for (int i = 0; i < K; i++) {
            double wt = c->w[i % Np];
            if (dim >= 2) wt *= c->w[(i / Np) % Np];
            if (dim == 3) wt *= c->w[i / (Np * Np)];
            c->wJ[(size_t)e * K + i] = wt * m.J[(size_t)e * K + i];
        }
    *out = c;
    return 0;
}

void orc_destroy(orc_ctx *c) { delete c; }

int orc_sizes(const orc_ctx *c, int *NE, int *K, int *nv)
{
    *NE = c->m.NE; *K = c->m.K; *nv = c->m.nv;
    return 0;
}

int orc_coords(const orc_ctx *c, double *X)
{
    const Mesh &m = c->m;
    const size_t NK = (size_t)m.NE * m.K;
    if (m.quad) {
        std::memcpy(X, m.X.data(), m.X.size() * sizeof(double));
        return 0;
    }
    for (int e = 0; e < m.NE; e++) {
        int idx[3] = {e % m.n[0], (e / m.n[0]) % m.n[1], e / (m.n[0] * m.n[1])};
        for (int i = 0; i < m.K; i++) {
            int ii[3] = {i % m.Np, (i / m.Np) % m.Np, i / (m.Np * m.Np)};
            for (int d = 0; d < m.dim; d++)
                X[d * NK + (size_t)e * m.K + i] = m.lo[d] + (idx[d] + c->x[ii[d]]) * m.h[d];
        }
    }
    return 0;
}

int orc_weights(const orc_ctx *c, double *WJ)
{
    std::memcpy(WJ, c->wJ.data(), c->wJ.size() * sizeof(double));
    return 0;
}

int orc_set_time(orc_ctx *c, double t)
{
    c->time = t;
    return 0;
}

double orc_get_time(const orc_ctx *c) { return c->time; }

int orc_set_boundary_split(orc_ctx *c, int attr, const double *normal, double offset, double speed,
                           const double *below, const double *above)
{
    if (!c->m.quad || attr < 1) return 1;
    const size_t na = c->m.attr_state.size() / 5;
    if ((size_t)attr > na) return 1;
    if (c->attr_fn.size() < 16 * na) c->attr_fn.assign(16 * na, 0.0);
    double *f = &c->attr_fn[16 * (size_t)(attr - 1)];
    f[0] = 1.0;
    for (int d = 0; d < 3; d++) f[1 + d] = normal[d];
    f[4] = offset;
    f[5] = speed;
    for (int v = 0; v < 5; v++) { f[6 + v] = below[v]; f[11 + v] = above[v]; }
    return 0;
}

int orc_rhs(orc_ctx *c, const double *u, double *dudt)
{
    c->t_stage = c->time;
    Engine<double> E(c);
    E.rhs(u, dudt, c->p.threads);
    if (c->p.parabolic) {
        const size_t n = (size_t)c->m.dim * c->m.nv * c->m.NE * c->m.K;
        c->grad.resize(n);
        double *dst = c->grad.data();
        const double *src = E.G.get();
        const int nth = c->p.threads > 0 ? c->p.threads : 1;
        (void)nth;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nth) schedule(static)
#endif
        for (long long i = 0; i < (long long)n; i++) dst[i] = src[i];
    }
    return 0;
}

int orc_rk3_step(orc_ctx *c, double *u, double dt)
{
    rk3_step(c, u, dt);
    return 0;
}

double orc_max_wave_speed(const orc_ctx *c) { return c->lam; }
void orc_reset_wave_speed(orc_ctx *c) { c->lam = -1.0; }
double orc_hmin(const orc_ctx *c) { return c->m.hmin; }

int orc_alpha(const orc_ctx *c, double *alpha)
{
    std::memcpy(alpha, c->alpha.data(), c->alpha.size() * sizeof(double));
    return 0;
}

int orc_eta(const orc_ctx *c, double *eta)
{
    std::memcpy(eta, c->eta.data(), c->eta.size() * sizeof(double));
    return 0;
}

int orc_axis_recover(const orc_ctx *c, const double *rU, double *U)
{
    if (!c->p.axisymmetric) return 1;
    axis_recover(c, rU, U);
    return 0;
}

int orc_last_gradients(const orc_ctx *c, double *grad)
{
    if (c->grad.empty()) return 1;
    std::memcpy(grad, c->grad.data(), c->grad.size() * sizeof(double));
    return 0;
}

int orc_rhs_ld(orc_ctx *c, const double *u, double *dudt)
{
    const size_t n = (size_t)c->m.nv * c->m.NE * c->m.K;
    std::vector<long double> U(n), dU(n);
    for (size_t i = 0; i < n; i++) U[i] = u[i];
    Engine<long double> E(c);
    E.rhs(U.data(), dU.data(), c->p.threads);
    for (size_t i = 0; i < n; i++) dudt[i] = (double)dU[i];
    return 0;
}

int orc_count_rhs_flops(orc_ctx *c, const double *u, double *out)
{
    const size_t n = (size_t)c->m.nv * c->m.NE * c->m.K;
    std::vector<Cnt> U(n), dU(n);
    for (size_t i = 0; i < n; i++) U[i] = Cnt(u[i]);
    for (int i = 0; i < 8; i++) g_ops[i] = 0;
    double lam = c->lam;
    std::vector<double> al = c->alpha, et = c->eta;
    Engine<Cnt> E(c);
    E.rhs(U.data(), dU.data(), 1);
    c->lam = lam; c->alpha = al; c->eta = et;
    long long tot = 0;
    for (int i = 1; i < 8; i++) tot += g_ops[i];
    out[0] = (double)tot;
    for (int i = 1; i < 8; i++) out[i] = (double)g_ops[i];
    return 0;
}

} /* extern "C" */
