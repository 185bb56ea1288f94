// DGSEM right-hand side and fused SSP-RK3 stage update for MI355X (gfx950).
//
// Two kernels per RHS evaluation:
//
//  k_visc (parabolic only)  u -> BR1 lifting of the entropy variables,
//      EntropyGrad2PrimGrad, viscous flux, contravariant transform and the
//      weak-form derivative sum D^ Fv of the volume term; writes the raw
//      viscous volume sum (nv-1 per node) and the face-normal component of
//      the viscous flux at every face node of the element (nv-1 per face
//      node).  Restates DGSEMOperator.cpp:82-159, LiftingBR1.cpp:10-272 and
//      DGSEMIntegrator.cpp:617-641,705-738 (and 435-549 for the face side).
//
//  k_rhs  the inviscid flux differencing, surface fluxes, modal indicator and
//      subcell FV blending, plus the viscous volume/face terms from k_visc,
//      then either writes the RHS (Operator::Mult) or the RK3 stage
//      combination.  Restates DGSEMIntegrator.cpp:151-433,435-549,601-864,
//      972-1196, PerssonPeraireIndicator.cpp:21-58, DGSEMOperator.cpp:60-80
//      and MFEM's RK3SSPSolver stage arithmetic.
//
// Work decomposition: a workgroup owns EPB whole elements (one node per
// thread in the node phases).  For each direction the element's two-point
// fluxes are "jobs" — symmetric volume pairs (computed once, as the
// reference does), surface points and subcell interfaces — computed by all
// threads into LDS, then reduced per node in the reference's summation
// order.  Face jobs are computed by both neighbouring elements ("owner
// computes") with the face's canonical orientation, so each face point's flux
// is bitwise the same on both sides and no atomics or scatter are needed.
// k_rhs is instantiated for DIM = 1, 2 and 3 (1-D: one line per element, a
// single point per face, unit face normals; Euler only, so no k_visc).
//
// Both kernels are also built for forced workgroup sizes (64, 128 or 256
// threads, "launch shapes"); prandtl_autotune picks one per context by
// timing them on the caller's state.  k_ens runs k_rhs's tile body over the
// element tiles of many independent 2-D cases in one launch per RK stage.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/prandtl_hip.h"
#include "dg_physics.h"
#include "mesh.h"

namespace prandtl {

constexpr int MAXNP = 8;

#ifndef PRANDTL_WAVES
#define PRANDTL_WAVES 3
#endif

struct Tables {
    double D[MAXNP * MAXNP];    // D_ij (row-major), DGSEMIntegrator.cpp:58-69
    double Dh[MAXNP * MAXNP];   // D + E            (lines 71-74, un-transposed)
    double Dh2[MAXNP * MAXNP];  // 2D + E           (lines 76-80, un-transposed)
    double V1i[MAXNP * MAXNP];  // inverse 1-D shifted-Legendre Vandermonde
    double fvs[MAXNP];          // 1/(J w_i)            (line 1046)
    double w[MAXNP];            // LGL weights on [0,1]
    double x[MAXNP];            // LGL nodes on [0,1]
    double subn[3][MAXNP];      // subcell normal magnitude per interface (1199-1304)
};

struct Geo {
    int n[3];
    int NE;
    double nrm[3];   // adj(r,r) = face measure of an r-face
    double invJ;     // 1/J
    double aface;    // 1/(w_0 J)
    int bcmask;      // boundary type (BC_*) of side (r, s) in bits 3(2r+s)..+2; 0: periodic box
    const double *bcs;    // device outer states of inflow / specified-state boundaries, 5 per
                          // box side [3][2] (Cartesian) or per attribute (unstructured meshes)
    // general geometry of quadrilateral and hexahedral meshes (null on
    // Cartesian boxes), local faces f = 2r + side; per-node tables hold DIM*DIM
    // adjugate entries, per-point normals DIM components
    const double *met;    // [4][NE*K] adjugate row r, component d at 2r+d
    const double *jac;    // [NE*K] Jacobian determinant
    const int *fface;     // [NE*4] bit 0: first element of the face; bits 1-3 boundary type; 4+: attribute-1
    const int *fnb;       // [NE*4*NFP] neighbour node e2*K + n2 of each interior face point
    const int *fnf;       // [NE*4] 2 (e2*4 + f2) + reversed: the neighbour's face of each interior face
    const double *fnor;   // [NE*4*NFP*2] normal outward of the face's first element
    const double *subn;   // [NE][2][NFP][NP-1][2] subcell interface normals
    const double *corner; // hexahedral meshes: [NE][8][3] element corners; the per-node metrics
                          // are recomputed from them (met and jac are not stored)
    // axisymmetric quadrilateral meshes (null otherwise)
    const double *rnode;  // [NE*K] r = y of each node
    const int *axm;       // [NE*K] -1 off the axis, else 2 d + (1: radial RHS zeroed), d the descriptor
    const int *axj;       // [nax][3] element-local nodes j1, j2 (ray fit), jn (floor fallback); -1 none
    const double *axr;    // [nax][3] their r
    double axis_floor;    // density floor of the axis reconstruction
};

struct KArgs {
    Tables tab;
    Geo geo;
    Phys phys;
    int blend, parabolic, ec, quirks;
    int mode;             // 0: out = rhs; 1: out = y + dt rhs; 2: out = c0 u0 + c1 (y + dt rhs)
    double dt, c0, c1;
    const double *dtp;    // if set, dt is read from device memory (graph replays)
    const double *y;      // the state (axisymmetric: the weighted state r U)
    const double *u0;
    double *out;
    const void *visc;     // [nv-1][NE][K]            (FP32 in the mixed form)
    const void *fvf;      // [NE][2 dim][nv-1][Nfp]   (FP32 in the mixed form)
    double *alpha;        // [NE]
    double *eta;          // [NE]
    double *lam;          // per-block running maxima [gridDim.x]
    const int *skip;      // if set and nonzero, the launch does nothing (a finished device loop)
};

struct VArgs {
    Tables tab;
    Geo geo;
    Phys phys;
    int quirks;
    const double *y;
    void *visc;
    void *fvf;
    double *grad;         // optional [dim][nv][NE][K]
    const int *skip;      // as KArgs::skip
};

constexpr int cmx(int a, int b) { return a > b ? a : b; }
constexpr int pow2_at_least(int x) { return x <= 1 ? 1 : 2 * pow2_at_least((x + 1) / 2); }
constexpr int r64(int x) { return ((x + 63) / 64) * 64; }
// NP^d: nodes per element (d = DIM) and per face (d = DIM - 1)
constexpr int npow(int NP, int d) { return d <= 0 ? 1 : NP * npow(NP, d - 1); }

// Launch shapes of both kernels: 0 natural, 1..3 forced blocks of 64, 128
// and 256 threads (autotuned per case).
constexpr int NSHAPES = PRANDTL_SHAPES;
constexpr int shape_fb(int shape) { return shape == 1 ? 64 : shape == 2 ? 128 : shape == 3 ? 256 : 0; }

// Launch shape of k_visc: one thread per node, BLOCK / K elements per
// workgroup.
constexpr int visc_block(int DIM, int NP, int FB)
{
    const int K = npow(NP, DIM);
    return FB ? FB : K > 256 ? r64(K) : 256;
}
constexpr bool visc_shape_ok(int DIM, int NP, int shape)
{
    if (shape == 0) return true;
    if (shape < 0 || shape >= NSHAPES) return false;
    const int K = npow(NP, DIM);
    const int fb = shape_fb(shape);
    return K <= fb && fb != visc_block(DIM, NP, 0);
}
template <int DIM, int NP, int FB = 0> struct VCfg {
    static constexpr int NV = DIM + 2;
    static constexpr int K = npow(NP, DIM);
    static constexpr int NFP = npow(NP, DIM - 1);
    static constexpr int BLOCK = visc_block(DIM, NP, FB);
    static constexpr int EPB = BLOCK / K;
    // waves per SIMD to ask the register allocator for
    static constexpr int WAVES = K > 256 ? 2 : BLOCK <= 128 ? 4 : PRANDTL_WAVES;
    static_assert(EPB >= 1, "one node per thread");
};

template <int DIM, int NP>
__device__ __forceinline__ int line_node(int r, int l, int m)
{
    if (DIM == 1) return m;
    if (DIM == 2) return r == 0 ? l * NP + m : m * NP + l;
    const int a = l % NP, b = l / NP;
    if (r == 0) return (b * NP + a) * NP + m;
    if (r == 1) return (b * NP + m) * NP + a;
    return (m * NP + b) * NP + a;
}

template <int DIM, int NP>
__device__ __forceinline__ int line_of(int r, int i, int j, int k)
{
    if (DIM == 1) return 0;
    if (DIM == 2) return r == 0 ? j : i;
    if (r == 0) return k * NP + j;
    if (r == 1) return k * NP + i;
    return j * NP + i;
}

template <int NP> __device__ __forceinline__ int pair_index(int a, int b)
{
    // a < b; pairs ordered (0,1),(0,2),...,(0,N),(1,2),...
    return a * NP - (a * (a + 1)) / 2 + (b - a - 1);
}

template <int NP> __device__ __forceinline__ void pair_ab(int p, int &a, int &b)
{
    a = 0;
    int rem = p;
    while (rem >= NP - 1 - a) { rem -= NP - 1 - a; a++; }
    b = a + 1 + rem;
}

__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    // blocks are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of element groups so face neighbours share its L2
    const int q = nb >> 3, rr = nb & 7;
    const int x = bid & 7, k = bid >> 3;
    return (x < rr) ? x * (q + 1) + k : rr * (q + 1) + (x - rr) * q + k;
}

__device__ __forceinline__ int neighbor(const Geo &g, int e, int r, int side)
{
    int idx[3] = {e % g.n[0], (e / g.n[0]) % g.n[1], e / (g.n[0] * g.n[1])};
    idx[r] = side ? (idx[r] + 1 == g.n[r] ? 0 : idx[r] + 1) : (idx[r] == 0 ? g.n[r] - 1 : idx[r] - 1);
    return (idx[2] * g.n[1] + idx[1]) * g.n[0] + idx[0];
}

// Boundary type of element e's face (r, side), BC_PERIODIC for a face with
// a neighbour.
__device__ __forceinline__ int boundary_of(const Geo &g, int e, int r, int side)
{
    const int idx = r == 0 ? e % g.n[0] : r == 1 ? (e / g.n[0]) % g.n[1] : e / (g.n[0] * g.n[1]);
    const int nr = r == 0 ? g.n[0] : r == 1 ? g.n[1] : g.n[2];
    const bool at = side ? idx == nr - 1 : idx == 0;
    return at ? (g.bcmask >> (3 * (2 * r + side))) & 7 : BC_PERIODIC;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (v < o) ? o : v;
    }
    return v;
}

// Running maximum wave speed: each block keeps its own slot (one thread, a
// plain read-modify-write, no contention); prandtl_max_wave_speed reduces
// the slots when asked.  old: the slot's value, read by thread 0 when the
// block started (only this block writes the slot during a launch).  Non-finite speeds are ignored, as std::max(current,
// NaN) does in the reference (DGSEMIntegrator.cpp:270).
template <int BLOCK>
__device__ __forceinline__ void block_lam_max(double lam, double old, double *slots, double *sred, int slot)
{
    lam = wave_max(lam);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sred[w] = lam;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sred[0];
        for (int i = 1; i < (BLOCK + 63) / 64; i++) m = (m < sred[i]) ? sred[i] : m;
        if (old < m) slots[slot] = m;
    }
}

// ---------------------------------------------------------------------------
// k_rhs
// ---------------------------------------------------------------------------
// Launch shape of k_rhs: per element the work items are K nodes, JF surface
// points (all 2*DIM faces) and, per direction, JP volume pairs plus JV
// subcell interfaces.  EPB elements per workgroup are chosen so that each
// phase gives every thread about one item (3-D N=3: 2 elements, 192
// threads: 96 surface points and 96 pairs per direction per element).
//
// FB (forced block size, 0 = natural) gives the other launch shapes the
// autotuner tries; rshape() computes a shape without instantiating RCfg, so
// the host can test which shapes exist.
struct RShape {
    int K, WORK, EPB, BLOCK;
};
constexpr RShape rshape(int DIM, int NP, bool BLEND, int FB)
{
    const int K = npow(NP, DIM);
    const int NFP = npow(NP, DIM - 1);
    const int JP = NFP * (NP * (NP - 1) / 2 + (BLEND ? NP - 1 : 0));
    const int WORK = cmx(K, cmx(JP, 2 * DIM * NFP));
    const int EPB = K > 128 ? 1 : cmx(1, (FB ? FB : 256) / WORK);
    const int BLOCK = FB ? FB : K > 256 ? r64(K) : (r64(EPB * WORK) > 512 ? 512 : r64(EPB * WORK));
    return RShape{K, WORK, EPB, BLOCK};
}
// A forced k_rhs shape exists when it holds one node per thread and differs
// from the natural one.
constexpr bool rhs_shape_ok(int DIM, int NP, bool BLEND, int shape)
{
    if (shape == 0) return true;
    if (shape < 0 || shape >= NSHAPES || DIM == 1) return false;
    const RShape n = rshape(DIM, NP, BLEND, 0);
    const RShape f = rshape(DIM, NP, BLEND, shape_fb(shape));
    return f.EPB * f.K <= f.BLOCK && !(f.EPB == n.EPB && f.BLOCK == n.BLOCK);
}

template <int DIM, int NP, bool BLEND, int FB = 0> struct RCfg {
    static constexpr int NV = DIM + 2;
    static constexpr int K = npow(NP, DIM);
    static constexpr int NFP = npow(NP, DIM - 1);
    static constexpr int NPAIR = NP * (NP - 1) / 2;
    static constexpr int NFV = BLEND ? NP - 1 : 0;
    static constexpr int JP = NFP * (NPAIR + NFV);
    static constexpr int JF = 2 * DIM * NFP;
    static constexpr int WORK = rshape(DIM, NP, BLEND, FB).WORK;
    static constexpr int EPB = rshape(DIM, NP, BLEND, FB).EPB;
    static constexpr int BLOCK = rshape(DIM, NP, BLEND, FB).BLOCK;
    static constexpr int lc_pick()
    {
        for (int lc = NFP; lc >= 1; lc--)
            if (NFP % lc == 0 && EPB * lc * (NPAIR + NFV) * NV * 8 <= 32 * 1024) return lc;
        return 1;
    }
    static constexpr int LC = lc_pick();
    static constexpr int FXN = LC * (NPAIR + NFV);
    // Registers for 4 waves per SIMD (128) where a workgroup has at most two
    // waves or more than 256 nodes: there, two workgroups per CU beat the
    // spills the cap causes (3-D N=7 with blending: 3.4 ms per launch at
    // 128 registers against 4.3 ms at 180).  Otherwise PRANDTL_WAVES.
    static constexpr int WAVES = (K > 256 || BLOCK <= 128) ? 4 : PRANDTL_WAVES;
    static_assert(EPB * K <= BLOCK, "one node per thread");
};

// LDS of one k_rhs tile (EPB elements).
template <int DIM, int NP, bool BLEND, bool QNS3, int FB, bool GEN = false, bool AXI = false> struct RhsSmem {
    using C = RCfg<DIM, NP, BLEND, FB>;
    static constexpr int NQ = DIM + 1 + (BLEND ? 2 : 0);   // v[DIM], beta, (1/beta, lambda)
    static constexpr int GK = QNS3 ? C::K : 1, GE = QNS3 ? C::EPB : 1;
    double Dh2[NP * NP];
    double V1i[BLEND ? NP * NP : 1];
    double Fvs[NP];
    double Sub[DIM][NP];
    double U[C::EPB][C::NV][C::K];
    double Q[C::EPB][NQ][C::K];
    // signed surface contributions (surface phase), then the pair and
    // subcell fluxes of one line chunk (direction phases): one buffer
    // (also the indicator's scratch: 4 K values per element)
    static constexpr int FF = cmx(cmx(C::JF, C::FXN), BLEND ? (4 * C::K + C::NV - 1) / C::NV : 1);
    static_assert(!BLEND || 4 * C::K <= C::NV * FF, "indicator scratch fits the flux buffer");
    double F[C::EPB][C::NV][FF];
    double G[GE][C::NV][GK];
    double Al[C::EPB];
    double Red[(C::BLOCK + 63) / 64];
    double M[GEN ? C::EPB : 1][GEN ? DIM * DIM : 1][GEN ? C::K : 1];   // per-node adjugate rows (general geometry)
    double R[AXI ? C::EPB : 1][AXI ? C::K : 1];      // per-node r (axisymmetric)
    double Cn[GEN && DIM == 3 ? C::EPB : 1][GEN && DIM == 3 ? 24 : 1];   // hexahedron corners
};

// DGSEMOperator::RecoverStateFromWeighted (DGSEMOperator.cpp:161-424) for one
// node q: U = rU / r off the axis; on it (low-order axis reconstruction, as
// an "axis" boundary selects) each of rho, rho u_z, rho E from the ray fit
// through the origin over the nodes j1, j2 of the node's descriptor,
// minmod-limited against 1.5 f_j / r_j and capped at 10 |f_j1| / r_j1 (one
// node: its f / r), radial momentum 0; a density below the floor, or a
// non-finite one or energy, takes node jn's state instead.  The nodes are
// chosen on the host from the geometry.
__device__ __forceinline__ double ax_minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    const double m = fmin(fabs(a), fabs(b));
    return a > 0.0 ? m : -m;
}

__device__ __forceinline__ double ax_clamp(double x, double m)
{
    if (!isfinite(x)) return 0.0;
    if (fabs(x) <= m) return x;
    return x > 0.0 ? m : (x < 0.0 ? -m : 0.0);
}

__device__ __forceinline__ void axis_node(const double *rU, const Geo &g, int K, size_t NK, size_t q, double U[4])
{
    const int code = g.axm[q];
    if (code < 0) {
        const double r = g.rnode[q];
#pragma unroll
        for (int v = 0; v < 4; v++) U[v] = rU[v * NK + q] / r;
        return;
    }
    const int d = code >> 1;
    const size_t eb = q - q % (size_t)K;
    const int j1 = g.axj[3 * d], j2 = g.axj[3 * d + 1], jn = g.axj[3 * d + 2];
    const double r1 = g.axr[3 * d], r2 = g.axr[3 * d + 1], rn = g.axr[3 * d + 2];
    auto fit = [&](int comp) -> double {
        if (j1 < 0) return 0.0;
        const double f1 = rU[comp * NK + eb + j1];
        const double fref = fabs(f1);
        const double cap = (fref < 1e-12) ? INFINITY : 10.0 * fref / fmax(r1, 1e-14);
        if (j2 >= 0) {
            const double f2 = rU[comp * NK + eb + j2];
            const double m = (r1 * f1 + r2 * f2) / fmax(r1 * r1 + r2 * r2, 1e-14);
            return ax_clamp(ax_minmod(m, ax_minmod(1.5 * (f1 / r1), 1.5 * (f2 / r2))), cap);
        }
        return ax_clamp(f1 / r1, cap);
    };
    const double rho = fit(0), mz = fit(1), E = fit(3);
    if (!isfinite(rho) || rho < g.axis_floor || !isfinite(E)) {
#pragma unroll
        for (int v = 0; v < 4; v++) U[v] = rU[v * NK + eb + jn] / rn;
        U[2] = 0.0;
    } else {
        U[0] = rho;
        U[1] = mz;
        U[2] = 0.0;
        U[3] = E;
    }
}

// Adjugate rows (row r = Ja^r, the contravariant metric) and determinant of
// a hexahedron's trilinear map at reference point xi, from its corners c
// ([8][3], MFEM's vertex order): the metrics the host builds per node, here
// recomputed where they are used, which saves the 80 bytes per node of a
// stored table.
__device__ __forceinline__ void hex_metrics(const double *c, double x0, double x1, double x2, double *adj,
                                            double &det)
{
    constexpr int CV[8] = {0, 1, 3, 2, 4, 5, 7, 6};   // vertex k -> bits (x + 2y + 4z) of its corner
    double J[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) J[a][b] = 0.0;
    const double xi[3] = {x0, x1, x2};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        double f[3], df[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const bool one = (CV[k] >> d) & 1;
            f[d] = one ? xi[d] : 1.0 - xi[d];
            df[d] = one ? 1.0 : -1.0;
        }
        const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) J[a][b] += dN[b] * c[3 * k + a];
    }
    adj[0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    adj[1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    adj[2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    adj[3] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    adj[4] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    adj[5] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    adj[6] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    adj[7] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    adj[8] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    det = J[0][0] * adj[0] + J[0][1] * adj[3] + J[0][2] * adj[6];
}

// Face point l (a + NP b, a and b the other two coordinates in increasing
// order) of a 3-D element's node on a face of direction r.
template <int NP> __device__ __forceinline__ int face_point_3(int node, int r)
{
    const int i = node % NP, j = (node / NP) % NP, k = node / (NP * NP);
    return r == 0 ? j + NP * k : (r == 1 ? i + NP * k : i + NP * j);
}

// One tile of k_rhs: elements [tile*EPB, tile*EPB + EPB) of the case A
// describes; slot indexes the tile's running wave-speed maximum.
// BLEND: modal indicator + subcell FV blending compiled in.
// QNS3: the reference's NS-3D zeta-partner defect (faithful NS in 3-D).
// VT: element type of the viscous inputs from k_visc (float in the mixed form).
// FB: forced workgroup size (0: the natural one of RCfg).
// OPEN: boundary sides on a Cartesian box.  GEN: general geometry of a
// quadrilateral or hexahedral mesh (per-node metrics, face tables; boundaries
// by attribute).  AXI
// (with GEN, Euler): axisymmetric, A.y the weighted state r U, recovered
// node by node as it is loaded (this element's nodes and the neighbours'
// face points); two-point, subcell and face fluxes weighted by r, the radial
// momentum's pressure source, its RHS zeroed on the axis.
template <int DIM, int NP, bool BLEND, bool QNS3, typename VT, int FB, bool OPEN = false, bool GEN = false,
          bool AXI = false>
__device__ __forceinline__ void rhs_body(const KArgs &A, const int tile, const int slot,
                                         RhsSmem<DIM, NP, BLEND, QNS3, FB, GEN, AXI> &S)
{
    using C = RCfg<DIM, NP, BLEND, FB>;
    constexpr int NV = C::NV, K = C::K, NFP = C::NFP, NPAIR = C::NPAIR, NFV = C::NFV;
    constexpr int EPB = C::EPB, LC = C::LC, BLOCK = C::BLOCK, FXN = C::FXN, JF = C::JF;
    constexpr int NQ = RhsSmem<DIM, NP, BLEND, QNS3, FB, GEN, AXI>::NQ;
    static_assert(!AXI || (GEN && DIM == 2), "axisymmetric meshes are quadrilateral meshes");
    constexpr int MC = DIM * DIM;   // adjugate entries per node (general geometry)
    auto &sM = S.M;
    (void)sM;
    auto &sR = S.R;
    (void)sR;
    (void)FXN;
    auto &sDh2 = S.Dh2;
    auto &sV1i = S.V1i;
    auto &sFvs = S.Fvs;
    auto &sSub = S.Sub;
    auto &sU = S.U;
    auto &sQ = S.Q;
    auto &sF = S.F;
    auto &sFc = S.F;
    auto &sG = S.G;
    auto &sAl = S.Al;
    auto &sRed = S.Red;

    if (A.skip && *A.skip) return;   // uniform: before any barrier
    const int t = threadIdx.x;
    for (int i = t; i < NP * NP; i += BLOCK) {
        sDh2[i] = A.tab.Dh2[(i / NP) * MAXNP + i % NP];
        if (BLEND) sV1i[i] = A.tab.V1i[(i / NP) * MAXNP + i % NP];
    }
    for (int i = t; i < NP; i += BLOCK) sFvs[i] = A.tab.fvs[i];
    for (int i = t; i < DIM * NP; i += BLOCK) sSub[i / NP][i % NP] = A.tab.subn[i / NP][i % NP];

    const Phys &P = A.phys;
    const int NE = A.geo.NE;
    const size_t NK = (size_t)NE * K;
    const int ebase = tile * EPB;
    const int el = t / K;
    const int n = t - el * K;
    const int e = ebase + el;
    const bool act = (el < EPB) && (e < NE);
    const bool quirk3d = A.quirks != 0;
    const int ii[3] = {n % NP, (n / NP) % NP, DIM == 3 ? n / (NP * NP) : 0};

    // ---- prologue: own node state, the neighbour data of this thread's
    // first surface point, the step size and the tile's wave-speed slot, all
    // loads in flight together ----
    const double dt = A.dtp ? *A.dtp : A.dt;
    const double lam_old = (t == 0) ? A.lam[slot] : 0.0;
    double ys[NV];
    if (act)
#pragma unroll
        for (int v = 0; v < NV; v++) ys[v] = A.y[v * NK + (size_t)e * K + n];
    // general geometry: this node's adjugate rows and Jacobian (and r)
    double jn = 1.0, mown[GEN ? MC : 1], rown = 0.0;
    int axown = -1;
    double yw[AXI ? NV : 1];   // axisymmetric: the weighted state of the update
#pragma unroll
    for (int c = 0; c < (GEN ? MC : 1); c++) mown[c] = 0.0;
    if (AXI && act) {
        rown = A.geo.rnode[(size_t)e * K + n];
        axown = A.geo.axm[(size_t)e * K + n];
#pragma unroll
        for (int v = 0; v < NV; v++) yw[v] = ys[v];
        axis_node(A.y, A.geo, K, NK, (size_t)e * K + n, ys);
    }
    if constexpr (GEN && DIM == 3) {
        // the tile's corners into LDS, then each node's metrics from them
        for (int i = t; i < EPB * 24; i += BLOCK)
            S.Cn[i / 24][i % 24] = (ebase + i / 24 < NE) ? A.geo.corner[(size_t)(ebase + i / 24) * 24 + i % 24] : 0.0;
        __syncthreads();
        if (act) hex_metrics(S.Cn[el], A.tab.x[ii[0]], A.tab.x[ii[1]], A.tab.x[ii[2]], mown, jn);
    } else if constexpr (GEN) {
        if (act) {
            jn = A.geo.jac[(size_t)e * K + n];
#pragma unroll
            for (int c = 0; c < MC; c++) mown[c] = A.geo.met[c * NK + (size_t)e * K + n];
        }
    }
    // Blended kernels carry the subcell-flux work and its accumulator; they
    // do not hold loads in flight across phases (the surface data is loaded
    // in the surface phase, the assembly inputs at the assembly), which keeps
    // their register count down.
    constexpr bool PREFETCH = !BLEND && !GEN;
    double pn[NV], pfo[NV], pfn[NV];
    if constexpr (PREFETCH) {
        const int jf = t;
        const int ej = jf / JF;
        const int eo = ebase + ej;
        if (jf < EPB * JF && eo < NE) {
            const int rem = jf - ej * JF;
            const int f = rem / NFP, l = rem - f * NFP;
            const int r = f >> 1, side = f & 1;
            const int en = neighbor(A.geo, eo, r, side);
            const int nn = line_node<DIM, NP>(r, l, side ? 0 : NP - 1);
#pragma unroll
            for (int v = 0; v < NV; v++) pn[v] = A.y[v * NK + (size_t)en * K + nn];
            if (A.parabolic) {
                const VT *fo = static_cast<const VT *>(A.fvf) + (((size_t)eo * 2 * DIM + f) * (NV - 1)) * NFP;
                const VT *fn = static_cast<const VT *>(A.fvf) + (((size_t)en * 2 * DIM + (f ^ 1)) * (NV - 1)) * NFP;
#pragma unroll
                for (int v = 1; v < NV; v++) { pfo[v] = fo[(v - 1) * NFP + l]; pfn[v] = fn[(v - 1) * NFP + l]; }
            }
        }
    }

    // ---- node phase: state and per-node flux quantities into LDS ----
    double lam = 0.0;
    if (act) {
#pragma unroll
        for (int v = 0; v < NV; v++) sU[el][v][n] = ys[v];
        if constexpr (GEN)
#pragma unroll
            for (int c = 0; c < MC; c++) sM[el][c][n] = mown[c];
        if constexpr (AXI) sR[el][n] = rown;
        const NodeQ<DIM> q = node_quants<DIM>(ys, P);
#pragma unroll
        for (int d = 0; d < DIM; d++) sQ[el][d][n] = q.v[d];
        sQ[el][DIM][n] = q.beta;
        if constexpr (BLEND) {
            sQ[el][DIM + 1][n] = q.ibeta;
            sQ[el][DIM + 2][n] = q.lam;
        }
        lam = (q.lam == q.lam) ? q.lam : 0.0;
        if (QNS3 && ii[1] == NP - 1)
#pragma unroll
            for (int v = 0; v < NV; v++) sG[el][v][n] = 0.0;
    }
    __syncthreads();

    // ---- modal indicator and blending coefficient (element-local) ----
    if (BLEND) {
        double *sX = &sF[0][0][0];
        double *sY = sX + EPB * K;
        if (act) {
            // rho * ComputePressure(state) with V /= rho (BasicOperations.cpp:31-37)
            const double rr = 1.0 / ys[0];
            double vv = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                const double v = ys[1 + d] * rr;
                vv = (d == 0) ? v * v : vv + v * v;
            }
            sX[el * K + n] = ys[0] * (P.gm1 * (ys[DIM + 1] - 0.5 * ys[0] * vv));
        }
        __syncthreads();
        // modes = V^-1 q, V^-1 = kron of the 1-D inverses, one direction at a time
        double *src = sX, *dst = sY;
#pragma unroll
        for (int r = 0; r < DIM; r++) {
            if (act) {
                const int a = ii[r];
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    int jj[3] = {ii[0], ii[1], ii[2]};
                    jj[r] = i;
                    const int m = DIM == 2 ? jj[1] * NP + jj[0] : (jj[2] * NP + jj[1]) * NP + jj[0];
                    s = (i == 0) ? sV1i[a * NP + i] * src[el * K + m] : s + sV1i[a * NP + i] * src[el * K + m];
                }
                dst[el * K + n] = s;
            }
            __syncthreads();
            double *tmp = src; src = dst; dst = tmp;
        }
        // modal energies: all modes, modes of degree <= N-1 and <= N-2 in
        // every direction (masks of PerssonPeraireIndicator.cpp:38-52).  Node
        // n holds the mode of degrees ii[] (mode order of ModalBasis.cpp:72-104
        // maps to this position); the three sums are a binary tree over the
        // element's K entries, the same for every launch shape.
        double *sE0 = dst, *sE1 = sX + 2 * EPB * K, *sE2 = sX + 3 * EPB * K;
        if (act) {
            const double md = src[el * K + n];
            bool in1 = true, in2 = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                if (ii[d] > NP - 3) in2 = false;
                if (ii[d] > NP - 2) in1 = false;
            }
            const double e2 = md * md;
            sE0[el * K + n] = e2;
            sE1[el * K + n] = in1 ? e2 : 0.0;
            sE2[el * K + n] = in2 ? e2 : 0.0;
        }
        __syncthreads();
        constexpr int HALF = pow2_at_least(K) / 2;
#pragma unroll
        for (int h = HALF; h >= 1; h >>= 1) {
            if (act && n < h && n + h < K) {
                const int i0 = el * K + n;
                sE0[i0] += sE0[i0 + h];
                sE1[i0] += sE1[i0 + h];
                sE2[i0] += sE2[i0 + h];
            }
            __syncthreads();
        }
        double mm = 0.0, m1 = 0.0, m2 = 0.0;
        if (act && n == 0) { mm = sE0[el * K]; m1 = sE1[el * K]; m2 = sE2[el * K]; }
        if (act && n == 0) {
            double eta = 1.0 - m1 / mm;
            const double eta2 = 1.0 - m2 / m1;
            eta = (eta < eta2) ? eta2 : eta;
            // DGSEMOperator.cpp:67-77, sharpness 9.21024 (DGSEMOperator.hpp:50)
            double al = 1.0 / (1.0 + exp(-9.21024 * (eta - P.thresh) / P.thresh));
            if (al < P.alpha_min) al = 0.0;
            else if (al > 1.0 - P.alpha_min) al = 1.0;
            al = (P.alpha_max < al) ? P.alpha_max : al;
            sAl[el] = al;
            A.alpha[e] = al;
            A.eta[e] = eta;
        }
        __syncthreads();
    }
    const double alpha_e = (BLEND && act) ? sAl[el] : 0.0;

    // ---- surface points of all faces (canonical orientation: elem1 = low
    // side, normal +e_r), signed contribution for this element ----
    for (int jf = t; jf < EPB * JF; jf += BLOCK) {
        const int ej = jf / JF;
        const int eo = ebase + ej;
        if (eo >= NE) continue;
        const int rem = jf - ej * JF;
        const int f = rem / NFP, l = rem - f * NFP;
        const int r = f >> 1, side = f & 1;
        const double nr = A.geo.nrm[r];
        if (!GEN && (!PREFETCH || jf != t)) {
            // not prefetched, or more surface points than threads: load now
            const int en = neighbor(A.geo, eo, r, side);
            const int nn = line_node<DIM, NP>(r, l, side ? 0 : NP - 1);
#pragma unroll
            for (int v = 0; v < NV; v++) pn[v] = A.y[v * NK + (size_t)en * K + nn];
            if (A.parabolic) {
                const VT *fo = static_cast<const VT *>(A.fvf) + (((size_t)eo * 2 * DIM + f) * (NV - 1)) * NFP;
                const VT *fn = static_cast<const VT *>(A.fvf) + (((size_t)en * 2 * DIM + (f ^ 1)) * (NV - 1)) * NFP;
#pragma unroll
                for (int v = 1; v < NV; v++) { pfo[v] = fo[(v - 1) * NFP + l]; pfn[v] = fn[(v - 1) * NFP + l]; }
            }
        }
        const int no = line_node<DIM, NP>(r, l, side ? NP - 1 : 0);
        double so[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) so[v] = sU[ej][v][no];
        if constexpr (GEN) {
            // general geometry: the face's normal and first element from the
            // face tables (owner computes, as on the box), boundary points by
            // attribute
            const size_t efo = (size_t)eo * (2 * DIM) + f;
            const int code = A.geo.fface[efo];
            const int bt = (code >> 1) & 7;
            double nvec[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) nvec[d] = A.geo.fnor[(efo * NFP + l) * DIM + d];
            // viscous traces Fv.n (the face's normal) of this side, and of the
            // neighbour's matching point
            const VT *fvfp = static_cast<const VT *>(A.fvf);
            double F[NV];
            if (bt != BC_PERIODIC) {
                const double *bs = A.geo.bcs + (code >> 4) * 5;
                const double lb = boundary_flux_vn<DIM>(bt, nvec, so, bs, P, quirk3d, F);
                lam = (lam < lb) ? lb : lam;
                if constexpr (AXI) {
                    // dU_face1 *= r (BdrFaceIntegrator.cpp:109-111)
                    const double rf = sR[ej][no];
#pragma unroll
                    for (int v = 0; v < NV; v++) F[v] *= rf;
                }
                double d[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) d[v] = -F[v];
                if (A.parabolic && (bt == BC_OUTFLOW || bt == BC_NOSLIP)) {
                    double fn[NV];
#pragma unroll
                    for (int v = 1; v < NV; v++) fn[v] = (double)fvfp[(efo * (NV - 1) + (v - 1)) * NFP + l];
                    if (bt == BC_NOSLIP) {
                        double en = bs[0] * fn[1], nn2 = nvec[0] * nvec[0];
#pragma unroll
                        for (int d = 1; d < DIM; d++) {
                            en = en + bs[d] * fn[1 + d];
                            nn2 = nn2 + nvec[d] * nvec[d];
                        }
                        fn[NV - 1] = en + bs[3] * sqrt(nn2);
                    }
#pragma unroll
                    for (int v = 1; v < NV; v++) d[v] += fn[v];
                }
#pragma unroll
                for (int v = 0; v < NV; v++) sFc[ej][v][rem] = d[v];
                continue;
            }
            const int gn = A.geo.fnb[efo * NFP + l];
            if constexpr (AXI) {
                axis_node(A.y, A.geo, K, NK, (size_t)gn, pn);
            } else {
#pragma unroll
                for (int v = 0; v < NV; v++) pn[v] = A.y[v * NK + (size_t)gn];
            }
            const bool own = code & 1;
            double *g1 = own ? so : pn;
            double *g2 = own ? pn : so;
            double fvn[NV];
            fvn[0] = 0.0;
            if (A.parabolic) {
                const int nf = A.geo.fnf[efo];
                const size_t fid = (size_t)(nf >> 1);
                int l2;
                if constexpr (DIM == 2) l2 = (nf & 1) ? NP - 1 - l : l;
                else l2 = face_point_3<NP>(gn - (int)(fid / 6) * K, (int)(fid % 6) >> 1);
#pragma unroll
                for (int v = 1; v < NV; v++)
                    fvn[v] = ((double)fvfp[(efo * (NV - 1) + (v - 1)) * NFP + l] +
                              (double)fvfp[(fid * (NV - 1) + (v - 1)) * NFP + l2]) * 0.5;
                if (A.quirks) {
                    // ComputeViscousFlux left velocities in the momentum slots
                    // (DGSEMIntegrator.cpp:471-481)
                    const double r1 = 1.0 / g1[0], r2 = 1.0 / g2[0];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { g1[1 + d] *= r1; g2[1 + d] *= r2; }
                }
            }
            const NodeQ<DIM> q1 = node_quants<DIM>(g1, P);
            const NodeQ<DIM> q2 = node_quants<DIM>(g2, P);
            two_point_flux_v<DIM, true>(nvec, norml2_v<DIM>(nvec), q1, q2, g1, g2, P, quirk3d, F);
            if constexpr (AXI) {
                // flux_num *= r (DGSEMIntegrator.cpp:190-205), the r of the
                // first element's node on both sides
                const double rf = own ? sR[ej][no] : A.geo.rnode[gn];
#pragma unroll
                for (int v = 0; v < NV; v++) F[v] *= rf;
            }
#pragma unroll
            for (int v = 0; v < NV; v++) {
                double d;
                if (A.parabolic) d = own ? (-F[v] + fvn[v]) : (F[v] - fvn[v]);
                else d = own ? -F[v] : F[v];
                sFc[ej][v][rem] = d;
            }
            continue;
        }
        if constexpr (OPEN) {
            // boundary point: -F_b for the outward normal at the interior
            // node (BdrFaceIntegrator.cpp:61-123, applied after the interior
            // faces by DGSEMNonlinearForm.cpp:776-826); its wave speed joins
            // the running maximum (DGSEMOperator.cpp:585-589)
            const int bt = boundary_of(A.geo, eo, r, side);
            if (bt != BC_PERIODIC) {
                double F[NV];
                const double ns = side ? nr : -nr;
                const double *bs = A.geo.bcs + (2 * r + side) * 5;
                const double lb = boundary_flux<DIM>(bt, r, ns, nr, so, bs, P, quirk3d, F);
                lam = (lam < lb) ? lb : lam;
                double d[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) d[v] = -F[v];
                if (A.parabolic && (bt == BC_OUTFLOW || bt == BC_NOSLIP)) {
                    // + Fv.n of the interior state and gradients (this
                    // element's trace from k_visc); inflow and specified
                    // states have zero gradients, so no viscous flux
                    // (BdrFaceIntegrator.cpp:141-175); the no-slip wall's energy
                    // term is V.(Fv.n)_mom + qn |n|
                    double fn[NV];
#pragma unroll
                    for (int v = 1; v < NV; v++) fn[v] = pfo[v] * ns;
                    if (bt == BC_NOSLIP) {
                        double en = bs[0] * fn[1];
#pragma unroll
                        for (int dd = 1; dd < DIM; dd++) en = en + bs[dd] * fn[1 + dd];
                        fn[NV - 1] = en + bs[3] * sqrt(ns * ns);
                    }
#pragma unroll
                    for (int v = 1; v < NV; v++) d[v] += fn[v];
                }
#pragma unroll
                for (int v = 0; v < NV; v++) sFc[ej][v][rem] = d[v];
                continue;
            }
        }
        double *s1 = side ? so : pn;
        double *s2 = side ? pn : so;
        double fvn[NV];
        fvn[0] = 0.0;
        if (A.parabolic) {
#pragma unroll
            for (int v = 1; v < NV; v++) fvn[v] = ((pfo[v] + pfn[v]) * 0.5) * nr;
            if (A.quirks) {
                // ComputeViscousFlux left velocities in the momentum slots of
                // state1/state2 (DGSEMIntegrator.cpp:471-481)
                const double r1 = 1.0 / s1[0], r2 = 1.0 / s2[0];
#pragma unroll
                for (int d = 0; d < DIM; d++) { s1[1 + d] *= r1; s2[1 + d] *= r2; }
            }
        }
        const NodeQ<DIM> q1 = node_quants<DIM>(s1, P);
        const NodeQ<DIM> q2 = node_quants<DIM>(s2, P);
        double F[NV];
        if (A.ec) two_point_flux<DIM, false>(r, nr, nr, q1, q2, nullptr, nullptr, P, false, F);
        else two_point_flux<DIM, true>(r, nr, nr, q1, q2, s1, s2, P, quirk3d, F);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            double d;
            if (A.parabolic) d = side ? (-F[v] + fvn[v]) : (F[v] - fvn[v]);
            else d = side ? -F[v] : F[v];
            sFc[ej][v][rem] = d;
        }
    }

    double accV[NV], accS[NV], accF[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) { accV[v] = 0.0; accS[v] = 0.0; accF[v] = 0.0; }
    double u0v[NV], vis[NV];

    // this node's surface terms, faces in direction order, low side first
    __syncthreads();
    const double aface = GEN ? 1.0 / (A.tab.w[0] * jn) : A.geo.aface;
    if (act) {
#pragma unroll
        for (int r = 0; r < DIM; r++) {
            const int l = line_of<DIM, NP>(r, ii[0], ii[1], ii[2]);
            if (ii[r] == 0)
#pragma unroll
                for (int v = 0; v < NV; v++) accS[v] += aface * sFc[el][v][(2 * r) * NFP + l];
            if (ii[r] == NP - 1)
#pragma unroll
                for (int v = 0; v < NV; v++) accS[v] += aface * sFc[el][v][(2 * r + 1) * NFP + l];
        }
    }
    __syncthreads();

    for (int r = 0; r < DIM; r++) {
        const double nr = A.geo.nrm[r];
        if (PREFETCH && r == DIM - 1 && act) {
            // inputs of the final assembly, in flight during the last direction
#pragma unroll
            for (int v = 0; v < NV; v++) {
                u0v[v] = (A.mode == 2) ? A.u0[v * NK + (size_t)e * K + n] : 0.0;
                vis[v] = (A.parabolic && v > 0)
                             ? (double)static_cast<const VT *>(A.visc)[(v - 1) * NK + (size_t)e * K + n] : 0.0;
            }
        }
        for (int l0 = 0; l0 < NFP; l0 += LC) {
            const int npj = EPB * LC * NPAIR, nvj = EPB * LC * NFV;
            for (int j = t; j < npj + nvj; j += BLOCK) {
                if (j < npj) {
                    // ---- symmetric volume pair (a < b) on line l ----
                    const int ej = j / (LC * NPAIR);
                    const int rem = j - ej * (LC * NPAIR);
                    const int lc = rem / NPAIR, p = rem - lc * NPAIR;
                    if (ebase + ej >= NE) continue;
                    int a, b;
                    pair_ab<NP>(p, a, b);
                    const int na = line_node<DIM, NP>(r, l0 + lc, a);
                    const int nb = line_node<DIM, NP>(r, l0 + lc, b);
                    NodeQ<DIM> qa, qb;
                    qa.rho = sU[ej][0][na]; qb.rho = sU[ej][0][nb];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { qa.v[d] = sQ[ej][d][na]; qb.v[d] = sQ[ej][d][nb]; }
                    qa.beta = sQ[ej][DIM][na]; qb.beta = sQ[ej][DIM][nb];
                    double F[NV];
                    if constexpr (GEN) {
                        // ComputeVolumeFlux with the mean of the two adjugate rows
                        double mm[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) mm[d] = (sM[ej][DIM * r + d][na] + sM[ej][DIM * r + d][nb]) * 0.5;
                        two_point_flux_v<DIM, false>(mm, 0.0, qa, qb, nullptr, nullptr, P, false, F);
                        if constexpr (AXI) {
                            // f *= r_hat, r at the pair's midpoint (DGSEMIntegrator.cpp:272-292)
                            const double rh = (sR[ej][na] + sR[ej][nb]) * 0.5;
#pragma unroll
                            for (int v = 0; v < NV; v++) F[v] *= rh;
                        }
                    } else {
                        two_point_flux<DIM, false>(r, nr, nr, qa, qb, nullptr, nullptr, P, false, F);
                    }
#pragma unroll
                    for (int v = 0; v < NV; v++) sF[ej][v][lc * NPAIR + p] = F[v];
                    if (QNS3 && r == 1 && b == NP - 1)
#pragma unroll
                        for (int v = 0; v < NV; v++) sG[ej][v][na] = F[v];
                } else if constexpr (BLEND) {
                    // ---- subcell FV interface i+1/2 on line l ----
                    const int jv = j - npj;
                    const int ej = jv / (LC * NFV);
                    const int rem = jv - ej * (LC * NFV);
                    const int lc = rem / NFV, i = rem - lc * NFV;
                    if (ebase + ej >= NE) continue;
                    const int slot = LC * NPAIR + lc * NFV + i;
                    if (!(sAl[ej] > 0.0)) {
#pragma unroll
                        for (int v = 0; v < NV; v++) sF[ej][v][slot] = 0.0;
                        continue;
                    }
                    const int na = line_node<DIM, NP>(r, l0 + lc, i);
                    const int nb = line_node<DIM, NP>(r, l0 + lc, i + 1);
                    double s1[NV], s2[NV];
                    NodeQ<DIM> qa, qb;
#pragma unroll
                    for (int v = 0; v < NV; v++) { s1[v] = sU[ej][v][na]; s2[v] = sU[ej][v][nb]; }
                    qa.rho = s1[0]; qb.rho = s2[0];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { qa.v[d] = sQ[ej][d][na]; qb.v[d] = sQ[ej][d][nb]; }
                    qa.beta = sQ[ej][DIM][na]; qb.beta = sQ[ej][DIM][nb];
                    qa.ibeta = sQ[ej][NQ - 2][na]; qb.ibeta = sQ[ej][NQ - 2][nb];
                    qa.lam = sQ[ej][NQ - 1][na]; qb.lam = sQ[ej][NQ - 1][nb];
                    double F[NV];
                    if constexpr (GEN) {
                        const double *sn = A.geo.subn + (((((size_t)(ebase + ej) * DIM + r) * NFP + l0 + lc) * (NP - 1) + i) * DIM);
                        double nv2[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) nv2[d] = sn[d];
                        two_point_flux_v<DIM, true>(nv2, norml2_v<DIM>(nv2), qa, qb, s1, s2, P, quirk3d, F);
                        if constexpr (AXI) {
                            // flux_num *= mid_r (DGSEMIntegrator.cpp:1030-1032, 1110-1112)
                            const double rh = (sR[ej][na] + sR[ej][nb]) * 0.5;
#pragma unroll
                            for (int v = 0; v < NV; v++) F[v] *= rh;
                        }
                    } else {
                        two_point_flux<DIM, true>(r, sSub[r][i], sSub[r][i], qa, qb, s1, s2, P, quirk3d, F);
                    }
#pragma unroll
                    for (int v = 0; v < NV; v++) sF[ej][v][slot] = F[v];
                }
            }
            __syncthreads();
            if (act) {
                const int l = line_of<DIM, NP>(r, ii[0], ii[1], ii[2]);
                if (l >= l0 && l < l0 + LC) {
                    const int lc = l - l0;
                    const int a = ii[r];
#pragma unroll
                    for (int m = 0; m < NP; m++) {
                        if (m == a) continue;
                        const double dm = sDh2[a * NP + m];
                        if (QNS3 && r == 2 && m < a) {
                            // H_inviscid(id2).SetCol(k, g) (DGSEMIntegrator.cpp:702)
                            const int nm = line_node<DIM, NP>(r, l, m);
#pragma unroll
                            for (int v = 0; v < NV; v++) accV[v] += dm * sG[el][v][nm];
                        } else {
                            const int p = m < a ? pair_index<NP>(m, a) : pair_index<NP>(a, m);
#pragma unroll
                            for (int v = 0; v < NV; v++) accV[v] += dm * sF[el][v][lc * NPAIR + p];
                        }
                    }
                    if (BLEND && alpha_e > 0.0) {
                        const int base = LC * NPAIR + lc * NFV;
                        // 1/(J_i w_i) (DGSEMIntegrator.cpp:1046) with this node's J
                        const double fvs = GEN ? 1.0 / (jn * A.tab.w[a]) : sFvs[a];
#pragma unroll
                        for (int v = 0; v < NV; v++) {
                            double val;
                            if (a == 0) val = (0.0 - sF[el][v][base + 0]) * fvs;
                            else if (a == NP - 1) val = sF[el][v][base + NP - 2] * fvs;
                            else val = (sF[el][v][base + a - 1] - sF[el][v][base + a]) * fvs;
                            accF[v] = (r == 0) ? val : accF[v] + val;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    // ---- assemble, stage update, store ----
    if (!PREFETCH && act) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            u0v[v] = (A.mode == 2) ? A.u0[v * NK + (size_t)e * K + n] : 0.0;
            vis[v] = (A.parabolic && v > 0)
                         ? (double)static_cast<const VT *>(A.visc)[(v - 1) * NK + (size_t)e * K + n] : 0.0;
        }
    }
    if (act) {
        double src = 0.0;
        if constexpr (AXI) {
            // PressureFromConservative (DGSEMIntegrator.cpp:1306-1315) of this node
            const double ke = 0.5 * ((sU[el][1][n] * sU[el][1][n] + sU[el][2][n] * sU[el][2][n]) / sU[el][0][n]);
            src = P.gm1 * (sU[el][3][n] - ke);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            double vol = -accV[v];
            if (BLEND) vol *= (1.0 - alpha_e);
            if (A.parabolic && v > 0) vol = vol + vis[v];
            vol *= GEN ? 1.0 / jn : A.geo.invJ;
            double tot = (BLEND && alpha_e > 0.0) ? accF[v] * alpha_e + vol : vol;
            if (AXI && v == 2) tot += src;   // DGSEMIntegrator.cpp:411-426
            tot += accS[v];
            // ZeroAxisRadialMom (DGSEMOperator.cpp:521-533)
            if (AXI && v == 2 && axown >= 0 && (axown & 1)) tot = 0.0;
            const size_t gi = v * NK + (size_t)e * K + n;
            const double yv = AXI ? yw[v] : sU[el][v][n];
            double o;
            if (A.mode == 0) o = tot;
            else if (A.mode == 1) o = yv + dt * tot;
            else o = A.c0 * u0v[v] + A.c1 * (yv + dt * tot);
            A.out[gi] = o;
        }
    }
    block_lam_max<BLOCK>(lam, lam_old, A.lam, sRed, slot);
}

// OPEN: the box has non-periodic sides (Euler, natural shape); the periodic
// kernels carry no boundary code.  GEN: a quadrilateral or hexahedral mesh.
template <int DIM, int NP, bool BLEND, bool QNS3, typename VT, int FB, bool OPEN = false, bool GEN = false,
          bool AXI = false>
__global__ void __launch_bounds__((RCfg<DIM, NP, BLEND, FB>::BLOCK))
__attribute__((amdgpu_waves_per_eu(RCfg<DIM, NP, BLEND, FB>::WAVES))) k_rhs(const KArgs A)
{
    __shared__ RhsSmem<DIM, NP, BLEND, QNS3, FB, GEN, AXI> S;
    rhs_body<DIM, NP, BLEND, QNS3, VT, FB, OPEN, GEN, AXI>(A, xcd_remap(blockIdx.x, gridDim.x), blockIdx.x, S);
}

// ---------------------------------------------------------------------------
// k_ens: one launch per RK stage over the element tiles of an ensemble of
// independent 2-D Euler cases with mixed polynomial degree and blending.
//
// Workgroup b takes tile b of a host-built list sorted by decreasing cost; the
// workgroup dispatcher hands workgroups to CUs in index order as slots free,
// so the list is the work queue and needs no device counter.  A persistent
// form (a loop over tiles inside the kernel, from an atomic counter or a
// grid stride) keeps the tile body inside a loop, and register allocation
// then spills about 1.3 KB per lane at 168 VGPRs; this form stays at the
// largest variant's own 107 VGPRs with no scratch.
// ---------------------------------------------------------------------------
struct EnsTile {
    int c;   // case
    int g;   // element group (tile) of that case
    int v;   // variant: 2*(NP-3) + blend
};

constexpr int ENS_BLOCK = 256;

template <int NP, bool BLEND> constexpr size_t ens_smem() { return sizeof(RhsSmem<2, NP, BLEND, false, ENS_BLOCK>); }
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t ENS_SMEM = cmax(cmax(cmax(ens_smem<3, false>(), ens_smem<3, true>()), cmax(ens_smem<4, false>(), ens_smem<4, true>())),
                                 cmax(cmax(ens_smem<5, false>(), ens_smem<5, true>()), cmax(ens_smem<6, false>(), ens_smem<6, true>())));

template <int NP, bool BLEND>
__device__ __forceinline__ void ens_tile(const KArgs &A, int g, double *smem)
{
    rhs_body<2, NP, BLEND, false, double, ENS_BLOCK>(A, g, g, *reinterpret_cast<RhsSmem<2, NP, BLEND, false, ENS_BLOCK> *>(smem));
}

// grid = number of tiles; every field of a tile is uniform over the workgroup
__global__ void __launch_bounds__(ENS_BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
k_ens(const KArgs *__restrict__ cases, const EnsTile *__restrict__ tiles)
{
    __shared__ __align__(16) double smem[(ENS_SMEM + 7) / 8];
    const EnsTile T = tiles[blockIdx.x];
    const KArgs &A = cases[T.c];
    switch (T.v) {
    case 0: ens_tile<3, false>(A, T.g, smem); break;
    case 1: ens_tile<3, true>(A, T.g, smem); break;
    case 2: ens_tile<4, false>(A, T.g, smem); break;
    case 3: ens_tile<4, true>(A, T.g, smem); break;
    case 4: ens_tile<5, false>(A, T.g, smem); break;
    case 5: ens_tile<5, true>(A, T.g, smem); break;
    case 6: ens_tile<6, false>(A, T.g, smem); break;
    case 7: ens_tile<6, true>(A, T.g, smem); break;
    default: break;
    }
}

constexpr int ENS_MAX_CASES = 256;
struct DtArgs { double dt[ENS_MAX_CASES]; };
__global__ void k_set_dts(double *dst, int n, const DtArgs a)
{
    const int i = threadIdx.x;
    if (i < n) dst[i] = a.dt[i];
}

// ---------------------------------------------------------------------------
// k_visc: entropy variables, BR1 lifting, gradient conversion, viscous fluxes
// ---------------------------------------------------------------------------
// MIXED: the mixed-precision form.  Entropy variables stay FP64; each
// element's are shifted by their value at node 0 (in FP64) and the BR1
// derivatives, gradient conversion, viscous flux and D^ sums then run in
// FP32 on those offsets (D annihilates constants, so the shift leaves the
// derivatives unchanged while keeping their significant digits).  The
// outputs visc and fvf are FP32 in that form.
// OPEN: a box with boundary sides; their face points take the boundary
// type's lifting flux (LiftingBR1.cpp:311-407) instead of the neighbour jump.
// GEN: a quadrilateral or hexahedral mesh (per-node metrics, face tables; traces are
// Fv.n with the face's normal).
template <int DIM, int NP, bool MIXED, int FB, bool OPEN = false, bool GEN = false>
__global__ void __launch_bounds__((VCfg<DIM, NP, FB>::BLOCK))
__attribute__((amdgpu_waves_per_eu(VCfg<DIM, NP, FB>::WAVES))) k_visc(const VArgs A)
{
    using C = VCfg<DIM, NP, FB>;
    constexpr int MC = DIM * DIM;   // adjugate entries per node (general geometry)
    using R = typename std::conditional<MIXED, float, double>::type;
    if (A.skip && *A.skip) return;   // uniform: before any barrier
    constexpr int NV = C::NV, K = C::K, NFP = C::NFP, EPB = C::EPB, BLOCK = C::BLOCK;
    // phase 1 (lifting) needs the entropy variables and neighbour traces,
    // phase 2 (viscous divergence) the contravariant fluxes: one LDS region
    constexpr int P1 = EPB * NV * K + EPB * 2 * DIM * NV * NFP;
    constexpr int P2 = EPB * DIM * (NV - 1) * K;
    constexpr int SCR = P1 > P2 ? P1 : P2;

    __shared__ R sD[NP * NP];
    __shared__ R sDh[NP * NP];
    __shared__ double sUl[EPB][NV];   // state of node K-1 (the faithful stale state)
    __shared__ double sCn[GEN && DIM == 3 ? EPB : 1][GEN && DIM == 3 ? 24 : 1];   // hexahedron corners
    __shared__ double sWref[EPB][NV];
    __shared__ R sScr[SCR];
    R (*sW)[NV][K] = reinterpret_cast<R (*)[NV][K]>(sScr);
    R (*sWN)[2 * DIM][NV][NFP] = reinterpret_cast<R (*)[2 * DIM][NV][NFP]>(sScr + EPB * NV * K);
    R (*sFV)[DIM][NV - 1][K] = reinterpret_cast<R (*)[DIM][NV - 1][K]>(sScr);

    const int t = threadIdx.x;
    for (int i = t; i < NP * NP; i += BLOCK) {
        sD[i] = (R)A.tab.D[(i / NP) * MAXNP + i % NP];
        sDh[i] = (R)A.tab.Dh[(i / NP) * MAXNP + i % NP];
    }
    const Phys &P = A.phys;
    const int NE = A.geo.NE;
    const size_t NK = (size_t)NE * K;
    const int ebase = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int el = t / K;
    const int n = t - el * K;
    const int e = ebase + el;
    const bool act = (el < EPB) && (e < NE);
    const int ii[3] = {n % NP, (n / NP) % NP, DIM == 3 ? n / (NP * NP) : 0};
    R *visc = reinterpret_cast<R *>(A.visc);
    R *fvf = reinterpret_cast<R *>(A.fvf);
    const R gm1 = (R)P.gm1, gm1i = (R)P.gm1i;
    const ViscCoef<R> vcoef{(R)P.mu, (R)P.lam_h, P.sutherland != 0, (R)P.R_gas, (R)P.gamma, (R)P.PrInv};

    // every global load of this thread in flight together: its own node and
    // the neighbour trace nodes of its face jobs
    constexpr int NJ = EPB * 2 * DIM * NFP;
    constexpr int JT = (NJ + BLOCK - 1) / BLOCK;
    double ys[NV], wv[NV];
    double tv[JT][NV];
    if (act)
#pragma unroll
        for (int v = 0; v < NV; v++) ys[v] = A.y[v * NK + (size_t)e * K + n];
    double jn = 1.0, mown[GEN ? MC : 1];
#pragma unroll
    for (int c = 0; c < (GEN ? MC : 1); c++) mown[c] = 0.0;
    if constexpr (GEN && DIM == 3) {
        for (int i = t; i < EPB * 24; i += BLOCK)
            sCn[i / 24][i % 24] = (ebase + i / 24 < NE) ? A.geo.corner[(size_t)(ebase + i / 24) * 24 + i % 24] : 0.0;
        __syncthreads();
        if (act) hex_metrics(sCn[el], A.tab.x[ii[0]], A.tab.x[ii[1]], A.tab.x[ii[2]], mown, jn);
    } else if constexpr (GEN) {
        if (act) {
            jn = A.geo.jac[(size_t)e * K + n];
#pragma unroll
            for (int c = 0; c < MC; c++) mown[c] = A.geo.met[c * NK + (size_t)e * K + n];
        }
    }
#pragma unroll
    for (int k = 0; k < JT; k++) {
        const int j = t + k * BLOCK;
        const int ej = j / (2 * DIM * NFP);
        const int eo = ebase + ej;
        if (j < NJ && eo < NE) {
            const int rem = j - ej * (2 * DIM * NFP);
            const int f = rem / NFP, q = rem - f * NFP;
            const int r = f >> 1, side = f & 1;
            int en = neighbor(A.geo, eo, r, side);
            int nn = line_node<DIM, NP>(r, q, side ? 0 : NP - 1);
            if (OPEN && boundary_of(A.geo, eo, r, side) != BC_PERIODIC) {
                // boundary point: this element's own face node
                en = eo;
                nn = line_node<DIM, NP>(r, q, side ? NP - 1 : 0);
            }
            if constexpr (GEN) {
                const size_t efo = (size_t)eo * (2 * DIM) + f;
                if ((A.geo.fface[efo] >> 1) & 7) {
                    en = eo;
                    nn = line_node<DIM, NP>(r, q, side ? NP - 1 : 0);
                } else {
                    const int gn = A.geo.fnb[efo * NFP + q];
                    en = gn / K;
                    nn = gn - en * K;
                }
            }
#pragma unroll
            for (int v = 0; v < NV; v++) tv[k][v] = A.y[v * NK + (size_t)en * K + nn];
        }
    }
    if (act) {
        if (n == K - 1)
#pragma unroll
            for (int v = 0; v < NV; v++) sUl[el][v] = ys[v];
        cons2entropy<DIM>(ys, P, wv);
        if (MIXED && n == 0)
#pragma unroll
            for (int v = 0; v < NV; v++) sWref[el][v] = wv[v];
    }
    if (MIXED) __syncthreads();
    if (act)
#pragma unroll
        for (int v = 0; v < NV; v++) sW[el][v][n] = MIXED ? (R)(wv[v] - sWref[el][v]) : (R)wv[v];
    // neighbour traces of the entropy variables
#pragma unroll
    for (int k = 0; k < JT; k++) {
        const int j = t + k * BLOCK;
        const int ej = j / (2 * DIM * NFP);
        const int eo = ebase + ej;
        if (j < NJ && eo < NE) {
            const int rem = j - ej * (2 * DIM * NFP);
            const int f = rem / NFP, q = rem - f * NFP;
            double w[NV];
            cons2entropy<DIM>(tv[k], P, w);
            const int code = GEN ? A.geo.fface[(size_t)eo * (2 * DIM) + f] : 0;
            const int bt = GEN ? ((code >> 1) & 7) : OPEN ? boundary_of(A.geo, eo, f >> 1, f & 1) : BC_PERIODIC;
            if ((OPEN || GEN) && bt != BC_PERIODIC) {
                // the lifting flux itself (not a trace) for the surface part
                double fb[NV];
                lift_flux<DIM>(bt, w, A.geo.bcs + (GEN ? (code >> 4) : f) * 5, P, A.quirks != 0, fb);
#pragma unroll
                for (int v = 0; v < NV; v++) sWN[ej][f][v][q] = (R)fb[v];
            } else {
#pragma unroll
                for (int v = 0; v < NV; v++) sWN[ej][f][v][q] = MIXED ? (R)(w[v] - sWref[ej][v]) : (R)w[v];
            }
        }
    }
    __syncthreads();

    R gp[DIM][NV];
    R yr[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) yr[v] = act ? (R)ys[v] : R(1);
    if constexpr (GEN) if (act) {
        // grad w = (sum_r (D-sum along r) adj row r) / J, then the face terms
        // f n / (w0 J) with each face point's normal (LiftingBR1.cpp:10-272)
        const R invJ = (R)(1.0 / jn), aface = (R)(1.0 / (A.tab.w[0] * jn)), half = 0.5;
        R sd[DIM][NV];
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            const int a = ii[c];
            const int l = line_of<DIM, NP>(c, ii[0], ii[1], ii[2]);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                R s = 0;
#pragma unroll
                for (int m = 0; m < NP; m++)
                    s = (m == 0) ? sW[el][v][line_node<DIM, NP>(c, l, m)] * sD[a * NP + m]
                                 : s + sW[el][v][line_node<DIM, NP>(c, l, m)] * sD[a * NP + m];
                sd[c][v] = s;
            }
        }
#pragma unroll
        for (int cc = 0; cc < DIM; cc++)
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                R a = sd[0][v] * (R)mown[cc];
#pragma unroll
                for (int c = 1; c < DIM; c++) a = a + sd[c][v] * (R)mown[DIM * c + cc];
                gp[cc][v] = a * invJ;
            }
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            const int a = ii[c];
            if (a == 0 || a == NP - 1) {
                const int side = (a == NP - 1), f = 2 * c + side;
                const int l = line_of<DIM, NP>(c, ii[0], ii[1], ii[2]);
                const size_t ef = (size_t)e * (2 * DIM) + f;
                const int code = A.geo.fface[ef];
                R nv2[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) nv2[d] = (R)A.geo.fnor[(ef * NFP + l) * DIM + d];
                const bool bnd = ((code >> 1) & 7) != BC_PERIODIC;
                const bool own = code & 1;
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    R fv;
                    if (bnd) fv = sWN[el][f][v][l];
                    else {
                        const R wo = sW[el][v][n], wn = sWN[el][f][v][l];
                        fv = own ? half * (wn - wo) : half * (wo - wn);
                    }
#pragma unroll
                    for (int cc = 0; cc < DIM; cc++) gp[cc][v] += aface * (fv * nv2[cc]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < DIM; c++) entgrad2primgrad<DIM, R>(yr, gm1, gm1i, A.quirks != 0, gp[c]);
        if (A.grad)
#pragma unroll
            for (int c = 0; c < DIM; c++)
#pragma unroll
                for (int v = 0; v < NV; v++) A.grad[(c * NV + v) * NK + (size_t)e * K + n] = (double)gp[c][v];
    }
    if (!GEN && act) {
        const R nrm[3] = {(R)A.geo.nrm[0], (R)A.geo.nrm[1], (R)A.geo.nrm[2]};
        const R invJ = (R)A.geo.invJ, aface = (R)A.geo.aface, half = 0.5;
        // volume part: grad_r w = sum_l D_{i_r l} w_l, times adj / J
        // (LiftingBR1.cpp:176-215)
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            const int a = ii[c];
            const int l = line_of<DIM, NP>(c, ii[0], ii[1], ii[2]);
#pragma unroll
            for (int v = 0; v < NV; v++) {
                R s = 0;
#pragma unroll
                for (int m = 0; m < NP; m++)
                    s = (m == 0) ? sW[el][v][line_node<DIM, NP>(c, l, m)] * sD[a * NP + m]
                                 : s + sW[el][v][line_node<DIM, NP>(c, l, m)] * sD[a * NP + m];
                gp[c][v] = (s * nrm[c]) * invJ;
            }
            // surface part: +1/2 (w2 - w1) n / (w0 J) on both sides (LiftingBR1.cpp:37-65)
            if (a == 0 || a == NP - 1) {
                const int side = (a == NP - 1);
                if (OPEN && boundary_of(A.geo, e, c, side) != BC_PERIODIC) {
                    // f n_c / (w0 J) with the outward normal (LiftingBR1.cpp:311-407)
                    const R nc = side ? nrm[c] : -nrm[c];
#pragma unroll
                    for (int v = 0; v < NV; v++) gp[c][v] += aface * (sWN[el][2 * c + side][v][l] * nc);
                } else {
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        const R wo = sW[el][v][n];
                        const R wn = sWN[el][2 * c + side][v][l];
                        const R f = side ? half * (wn - wo) : half * (wo - wn);
                        gp[c][v] += aface * (f * nrm[c]);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < DIM; c++) entgrad2primgrad<DIM, R>(yr, gm1, gm1i, A.quirks != 0, gp[c]);
        if (A.grad)
#pragma unroll
            for (int c = 0; c < DIM; c++)
#pragma unroll
                for (int v = 0; v < NV; v++) A.grad[(c * NV + v) * NK + (size_t)e * K + n] = (double)gp[c][v];
    }
    __syncthreads();   // sW/sWN are dead from here; sFV reuses their storage

    if (act) {
        R fl[DIM + 1][DIM];
        // volume viscous flux uses the stale node K-1 state (DGSEMIntegrator.cpp:620)
        R sv[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) sv[v] = A.quirks ? (R)sUl[el][v] : yr[v];
        viscous_flux<DIM, R>(sv, gp, gm1, gm1i, vcoef, fl);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int v = 1; v < NV; v++) {
                if constexpr (GEN) {
                    // contravariant: adj row r times the physical flux
                    R a = (R)mown[DIM * r] * fl[v - 1][0];
#pragma unroll
                    for (int c = 1; c < DIM; c++) a += (R)mown[DIM * r + c] * fl[v - 1][c];
                    sFV[el][r][v - 1][n] = a;
                } else {
                    sFV[el][r][v - 1][n] = (R)A.geo.nrm[r] * fl[v - 1][r];
                }
            }
        bool onface = false;
#pragma unroll
        for (int c = 0; c < DIM; c++) onface |= (ii[c] == 0 || ii[c] == NP - 1);
        if (onface) {
            if (A.quirks) viscous_flux<DIM, R>(yr, gp, gm1, gm1i, vcoef, fl);
#pragma unroll
            for (int c = 0; c < DIM; c++) {
                if (ii[c] == 0 || ii[c] == NP - 1) {
                    const int side = (ii[c] == NP - 1);
                    const int l = line_of<DIM, NP>(c, ii[0], ii[1], ii[2]);
                    R *dst = fvf + (((size_t)e * 2 * DIM + 2 * c + side) * (NV - 1)) * NFP;
                    if constexpr (GEN) {
                        // Fv.n with the face's normal (outward of its first element)
                        const size_t ef = (size_t)e * (2 * DIM) + 2 * c + side;
                        R nf[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) nf[d] = (R)A.geo.fnor[(ef * NFP + l) * DIM + d];
#pragma unroll
                        for (int v = 1; v < NV; v++) {
                            R a = fl[v - 1][0] * nf[0];
#pragma unroll
                            for (int d = 1; d < DIM; d++) a = a + fl[v - 1][d] * nf[d];
                            dst[(v - 1) * NFP + l] = a;
                        }
                    } else {
#pragma unroll
                        for (int v = 1; v < NV; v++) dst[(v - 1) * NFP + l] = fl[v - 1][c];
                    }
                }
            }
        }
    }
    __syncthreads();
    if (act) {
        // sum_r sum_l Dhat_{i_r l} (adj Fv)^r_l  (DGSEMIntegrator.cpp:705-738)
        R acc[NV - 1];
#pragma unroll
        for (int v = 0; v < NV - 1; v++) acc[v] = 0;
#pragma unroll
        for (int r = 0; r < DIM; r++) {
            const int a = ii[r];
            const int l = line_of<DIM, NP>(r, ii[0], ii[1], ii[2]);
#pragma unroll
            for (int m = 0; m < NP; m++) {
                const int nm = line_node<DIM, NP>(r, l, m);
                const R dm = sDh[a * NP + m];
#pragma unroll
                for (int v = 0; v < NV - 1; v++) acc[v] += sFV[el][r][v][nm] * dm;
            }
        }
#pragma unroll
        for (int v = 0; v < NV - 1; v++) visc[v * NK + (size_t)e * K + n] = acc[v];
    }
}

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
__global__ void k_fill(double *p, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// max over the per-block slots into out[0] (one block)
__global__ void k_max_reduce(const double *slots, int n, double *out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = (m < slots[i]) ? slots[i] : m;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int i = 1; i < 4; i++) r = (r < red[i]) ? red[i] : r;
        out[0] = r;
    }
}
__global__ void k_set_double(double *p, double v) { *p = v; }
__global__ void k_reset_int(int *p) { *p = 0; }

// State of a time loop run on the device (Simulation::Run's stepping,
// Simulation.cpp:893-907, 1007-1022): time, step size, the clipped step of
// the current step, done flag and step count.
struct LoopState {
    double t, dt, dt_real;
    int done, steps, bad;
};

// maximum of slots[0..n) over a block of LOOP_BLOCK threads (every thread
// gets it)
constexpr int LOOP_BLOCK = 1024;
__device__ __forceinline__ double block_max_slots(const double *slots, int n)
{
    __shared__ double red[LOOP_BLOCK / 64];
    double m = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < n; i += LOOP_BLOCK) m = (m < slots[i]) ? slots[i] : m;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    double r = red[0];
    for (int i = 1; i < LOOP_BLOCK / 64; i++) r = (r < red[i]) ? red[i] : r;
    return r;
}

// the step actually taken: min(dt, t_final - t)
__device__ __forceinline__ double clipped(double dt, double t, double t_final)
{
    const double r = t_final - t;
    return (dt < r) ? dt : r;
}

// first step of a variable-dt loop: dt = (cfl * hmin) / (lambda * (N+1)^2)
// from the running maximum wave speed, and its clipped step
__global__ void __launch_bounds__(LOOP_BLOCK) k_loop_dt(LoopState *L, const double *slots, int n, double cfl_hmin, double np1sq,
                          double t_final)
{
    const double lam = block_max_slots(slots, n);
    if (threadIdx.x == 0) {
        L->dt = cfl_hmin / (lam * np1sq);
        L->dt_real = clipped(L->dt, L->t, t_final);
    }
}

// after a step: t += dt_real; variable dt: dt from the running maximum;
// done when t >= t_final - 1e-8 dt or max_steps (if > 0) steps were taken;
// otherwise the next step's clipped dt
__global__ void __launch_bounds__(LOOP_BLOCK) k_loop_end(LoopState *L, const double *slots, int n, double cfl_hmin, double np1sq,
                           double t_final, int variable, int max_steps)
{
    if (L->done) return;   // uniform
    const double lam = variable ? block_max_slots(slots, n) : 0.0;
    if (threadIdx.x == 0) {
        const double t = L->t + L->dt_real;
        const int steps = L->steps + 1;
        const double dt = variable ? cfl_hmin / (lam * np1sq) : L->dt;
        const int done = (t >= t_final - 1e-8 * dt) || (max_steps > 0 && steps >= max_steps);
        L->t = t;
        L->steps = steps;
        L->dt = dt;
        if (!done) L->dt_real = clipped(dt, t, t_final);
        L->done = done;
    }
}

template <int DIM>
__global__ void k_check(const double *u, size_t NK, double gm1, int *flag)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NK) return;
    double s[DIM + 2];
    bool bad = false;
    for (int v = 0; v < DIM + 2; v++) {
        s[v] = u[v * NK + i];
        bad |= !isfinite(s[v]);
    }
    double vv = 0.0;
    for (int d = 0; d < DIM; d++) vv += (s[1 + d] / s[0]) * (s[1 + d] / s[0]);
    const double p = gm1 * (s[DIM + 1] - 0.5 * s[0] * vv);
    bad |= !(s[0] > 0.0) || !(p > 0.0);
    if (bad) atomicOr(flag, 1);
}

// Primitive output fields of Simulation.cpp:1046-1061: rho, velocity
// components mom_d/rho, and p = (gamma-1)(E - rho |v|^2 / 2), with |v|^2
// summed in the reference's order from the divided velocities.
template <int DIM>
__global__ void k_primitives(const double *u, double *prim, size_t NK, double gm1)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NK) return;
    const double rho = u[i];
    double vsq = 0.0;
    prim[i] = rho;
    for (int d = 0; d < DIM; d++) {
        const double v = u[(1 + d) * NK + i] / rho;
        prim[(1 + d) * NK + i] = v;
        vsq = (d == 0) ? v * v : vsq + v * v;
    }
    prim[(DIM + 1) * NK + i] = gm1 * (u[(DIM + 1) * NK + i] - 0.5 * rho * vsq);
}

// RecoverStateFromWeighted for every node (prandtl_axis_recover, state checks
// and output of axisymmetric contexts).  skip as KArgs::skip.
__global__ void __launch_bounds__(256) k_axis(const double *rU, double *U, Geo g, int K, const int *skip)
{
    if (skip && *skip) return;
    const size_t NK = (size_t)g.NE * K;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NK) return;
    double u[4];
    axis_node(rU, g, K, NK, q, u);
#pragma unroll
    for (int v = 0; v < 4; v++) U[v * NK + q] = u[v];
}

// The NaN/Inf check of Simulation.cpp:1025-1035 inside a device loop, after
// every step whose count is a multiple of `every`: a non-finite or
// non-positive density or pressure stops the loop (bad = done = 1) before
// the next step.  Every writer stores the same values.
template <int DIM>
__global__ void k_loop_check(LoopState *L, const double *u, size_t NK, double gm1, int every)
{
    if (L->done || L->steps % every) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NK) return;
    double s[DIM + 2];
    bool bad = false;
    for (int v = 0; v < DIM + 2; v++) {
        s[v] = u[v * NK + i];
        bad |= !isfinite(s[v]);
    }
    double vv = 0.0;
    for (int d = 0; d < DIM; d++) vv += (s[1 + d] / s[0]) * (s[1 + d] / s[0]);
    const double p = gm1 * (s[DIM + 1] - 0.5 * s[0] * vv);
    bad |= !(s[0] > 0.0) || !(p > 0.0);
    if (bad) {
        L->bad = 1;
        L->done = 1;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace host {

// LGL nodes/weights on [0,1] (the rule of DGSEMIntegrator.cpp:18-22):
// Newton on P_N' in long double from Chebyshev-Lobatto guesses.
void lgl(int N, std::vector<double> &x, std::vector<double> &w)
{
    const int Np = N + 1;
    std::vector<long double> z(Np), wz(Np);
    for (int i = 0; i < Np; i++) {
        long double t = (i == 0) ? -1.0L : (i == N ? 1.0L : -std::cos((long double)M_PI * i / N));
        for (int it = 0; it < 100 && i > 0 && i < N; it++) {
            long double p0 = 1.0L, p1 = t, d0 = 0.0L, d1 = 1.0L;
            for (int k = 1; k < N; k++) {
                long double p2 = ((2 * k + 1) * t * p1 - k * p0) / (k + 1);
                long double d2 = d0 + (2 * k + 1) * p1;
                p0 = p1; p1 = p2; d0 = d1; d1 = d2;
            }
            const long double PN = (N == 1) ? t : p1, dPN = (N == 1) ? 1.0L : d1;
            const long double dd = (2.0L * t * dPN - N * (N + 1) * PN) / (1.0L - t * t);
            const long double step = dPN / dd;
            t -= step;
            if (std::fabs((double)step) < 1e-22) break;
        }
        long double p0 = 1.0L, p1 = t;
        for (int k = 1; k < N; k++) {
            long double p2 = ((2 * k + 1) * t * p1 - k * p0) / (k + 1);
            p0 = p1; p1 = p2;
        }
        const long double PN = (N == 1) ? t : p1;
        z[i] = t;
        wz[i] = 2.0L / (N * (N + 1) * PN * PN);
    }
    x.resize(Np);
    w.resize(Np);
    for (int i = 0; i < Np; i++) {
        x[i] = (double)(0.5L * (1.0L + 0.5L * (z[i] - z[N - i])));
        w[i] = (double)(0.25L * (wz[i] + wz[N - i]));
    }
    x[0] = 0.0;
    x[N] = 1.0;
    if (Np % 2 == 1) x[N / 2] = 0.5;
}

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
        const long double d = a[c * 2 * n + c];
        for (int j = 0; j < 2 * n; j++) a[c * 2 * n + j] /= d;
        for (int r = 0; r < n; r++)
            if (r != c) {
                const long double f = a[r * 2 * n + c];
                for (int j = 0; j < 2 * n; j++) a[r * 2 * n + j] -= f * a[c * 2 * n + j];
            }
    }
    Ai.assign(n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Ai[i * n + j] = (double)a[i * 2 * n + n + j];
    return true;
}

} // namespace host
} // namespace prandtl

using namespace prandtl;

struct prandtl_ctx {
    int dim, order, Np, K, NE, nv, Nfp;
    prandtl_mesh_desc mesh;
    prandtl_params prm;
    Tables tab;
    Geo geo;
    Phys phys;
    double *d_visc = nullptr, *d_fvf = nullptr, *d_alpha = nullptr, *d_eta = nullptr;
    double *d_ya = nullptr, *d_yb = nullptr;
    double *d_lam = nullptr;       // per-block running maxima (NE slots)
    double *d_lam_out = nullptr;
    int *d_flag = nullptr;
    std::string err;
    // kernel timing (prandtl_profile_begin/end): events bracketing every launch
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;   // 0 = k_rhs, 1 = k_visc
    size_t ev_used = 0;
    // SSP-RK3 step captured as a hipGraph (dt read from d_dt), replayed per step
    bool use_graph = true;
    double *d_dt = nullptr;
    double *d_bcs = nullptr;       // [3][2][5] outer boundary states (open boxes only)
    void *d_geo = nullptr;         // general quadrilateral geometry tables
    bool gen = false;              // quadrilateral or hexahedral mesh (general-geometry kernels)
    bool axi = false;              // axisymmetric (state r U)
    double *d_axU = nullptr;       // [nv][NE][K] the state recovered from r U (checks, output)
    double hmin = 0.0;             // smallest element size (the dt policy's h_min)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const double *graph_u = nullptr;
    hipStream_t graph_stream = nullptr;
    const double *dtp = nullptr;  // set while launching stages for a graph
    double graph_dt = 0.0;        // value last copied to d_dt
    bool graph_dt_set = false;
    int rhs_shape = 0;            // k_rhs launch shape (only shapes rhs_shape_ok accepts are set)
    int visc_shape = 0;           // k_visc launch shape (visc_shape_ok)
    // device time loop (prandtl_run)
    LoopState *d_loop = nullptr;
    const int *skip = nullptr;    // set while launching the loop's stages
};

namespace {

std::string g_create_err;

int fail(prandtl_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

int hip_check(prandtl_ctx *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return PRANDTL_OK;
    return fail(c, PRANDTL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Event pair around one launch when profiling is on (start before, stop after).
struct ProfScope {
    prandtl_ctx *c;
    hipStream_t s;
    size_t i;
    ProfScope(prandtl_ctx *cc, hipStream_t ss, int kind) : c(cc), s(ss), i(0)
    {
        if (!c->prof) return;
        i = c->ev_used;
        while (c->ev_pool.size() < i + 2) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) { c->prof = false; return; }
            c->ev_pool.push_back(e);
        }
        c->ev_kind.resize((i + 2) / 2);
        c->ev_kind[i / 2] = kind;
        c->ev_used = i + 2;
        (void)hipEventRecord(c->ev_pool[i], s);
    }
    ~ProfScope()
    {
        if (c->prof && c->ev_used >= i + 2) (void)hipEventRecord(c->ev_pool[i + 1], s);
    }
};

template <int DIM, int NP, int FB> void launch_visc_fb(prandtl_ctx *c, const VArgs &a, hipStream_t s)
{
    using Cf = VCfg<DIM, NP, FB>;
    const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
    if (c->prm.precision == PRANDTL_MIXED)
        hipLaunchKernelGGL((k_visc<DIM, NP, true, FB>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
    else hipLaunchKernelGGL((k_visc<DIM, NP, false, FB>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
}

// Orders the general-geometry kernels are built for: 2-D N = 1..5, 3-D N = 2..4.
constexpr bool gen_built(int DIM, int NP) { return DIM == 2 ? NP <= 6 : DIM == 3 && NP <= 5; }

template <int DIM, int NP> int launch_visc_nd(prandtl_ctx *c, const double *y, double *grad, hipStream_t s)
{
    VArgs a;
    a.tab = c->tab; a.geo = c->geo; a.phys = c->phys;
    a.quirks = c->prm.quirks;
    a.y = y; a.visc = c->d_visc; a.fvf = c->d_fvf; a.grad = grad;
    a.skip = c->skip;
    ProfScope ps(c, s, 1);
    if constexpr (gen_built(DIM, NP)) {
        if (c->gen) {
            // quadrilateral or hexahedral mesh, natural shape
            using Cf = VCfg<DIM, NP, 0>;
            const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
            if (c->prm.precision == PRANDTL_MIXED)
                hipLaunchKernelGGL((k_visc<DIM, NP, true, 0, false, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
            else hipLaunchKernelGGL((k_visc<DIM, NP, false, 0, false, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
            return hip_check(c, hipGetLastError(), "k_visc launch");
        }
    }
    if (c->geo.bcmask) {
        // boundary sides: the open variant, natural shape
        using Cf = VCfg<DIM, NP, 0>;
        const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
        if (c->prm.precision == PRANDTL_MIXED)
            hipLaunchKernelGGL((k_visc<DIM, NP, true, 0, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_visc<DIM, NP, false, 0, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
        return hip_check(c, hipGetLastError(), "k_visc launch");
    }
    switch (c->visc_shape) {
    case 1:
        if constexpr (visc_shape_ok(DIM, NP, 1)) launch_visc_fb<DIM, NP, 64>(c, a, s);
        break;
    case 2:
        if constexpr (visc_shape_ok(DIM, NP, 2)) launch_visc_fb<DIM, NP, 128>(c, a, s);
        break;
    case 3:
        if constexpr (visc_shape_ok(DIM, NP, 3)) launch_visc_fb<DIM, NP, 256>(c, a, s);
        break;
    default:
        launch_visc_fb<DIM, NP, 0>(c, a, s);
        break;
    }
    return hip_check(c, hipGetLastError(), "k_visc launch");
}

// 1-D operators are Euler operators (prandtl_create refuses parabolic ones):
// no viscous kernel is built for them.
template <int DIM, int NP> int launch_visc(prandtl_ctx *c, const double *y, double *grad, hipStream_t s)
{
    if constexpr (DIM == 1) return fail(c, PRANDTL_ERR_UNSUPPORTED, "no viscous kernel in 1-D");
    else return launch_visc_nd<DIM, NP>(c, y, grad, s);
}

template <int DIM, int NP, bool BLEND, int FB>
void launch_rhs_fb(prandtl_ctx *c, const KArgs &a, hipStream_t s)
{
    using Cf = RCfg<DIM, NP, BLEND, FB>;
    const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
    if (c->prm.parabolic && c->prm.precision == PRANDTL_MIXED)
        hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, float, FB>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
    else hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, double, FB>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
}

template <int DIM, int NP, bool BLEND>
int launch_rhs_t(prandtl_ctx *c, const KArgs &a, hipStream_t s)
{
    ProfScope ps(c, s, 0);
    const bool qns3 = DIM == 3 && c->prm.parabolic && c->prm.quirks;
    if constexpr (gen_built(DIM, NP)) {
        if (c->gen) {
            // quadrilateral or hexahedral mesh, natural shape
            using Cf = RCfg<DIM, NP, BLEND>;
            const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
            if constexpr (DIM == 2) {
                if (c->axi) {
                    // the axisymmetric RHS (Euler), recovering U from r U as it loads
                    hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, double, 0, false, true, true>), dim3(nb),
                                       dim3(Cf::BLOCK), 0, s, a);
                    return hip_check(c, hipGetLastError(), "k_rhs launch");
                }
            } else {
                if (qns3) {
                    // the faithful NS-3D form on a hexahedral mesh
                    hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, true, double, 0, false, true>), dim3(nb),
                                       dim3(Cf::BLOCK), 0, s, a);
                    return hip_check(c, hipGetLastError(), "k_rhs launch");
                }
            }
            if (c->prm.parabolic && c->prm.precision == PRANDTL_MIXED)
                hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, float, 0, false, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
            else hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, double, 0, false, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
            return hip_check(c, hipGetLastError(), "k_rhs launch");
        }
    }
    if (c->geo.bcmask) {
        // boundary sides, natural shape
        using Cf = RCfg<DIM, NP, BLEND>;
        const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
        if constexpr (DIM == 3) {
            if (qns3) {
                // the faithful NS-3D form with boundary sides
                hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, true, double, 0, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
                return hip_check(c, hipGetLastError(), "k_rhs launch");
            }
        }
        if (c->prm.parabolic && c->prm.precision == PRANDTL_MIXED)
            hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, float, 0, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, false, double, 0, true>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
        return hip_check(c, hipGetLastError(), "k_rhs launch");
    }
    if (qns3) {
        // the faithful NS-3D form runs in the natural shape only
        using Cf = RCfg<DIM, NP, BLEND>;
        const int nb = (c->NE + Cf::EPB - 1) / Cf::EPB;
        hipLaunchKernelGGL((k_rhs<DIM, NP, BLEND, DIM == 3, double, 0>), dim3(nb), dim3(Cf::BLOCK), 0, s, a);
        return hip_check(c, hipGetLastError(), "k_rhs launch");
    }
    switch (c->rhs_shape) {
    case 1:
        if constexpr (rhs_shape_ok(DIM, NP, BLEND, 1)) launch_rhs_fb<DIM, NP, BLEND, 64>(c, a, s);
        break;
    case 2:
        if constexpr (rhs_shape_ok(DIM, NP, BLEND, 2)) launch_rhs_fb<DIM, NP, BLEND, 128>(c, a, s);
        break;
    case 3:
        if constexpr (rhs_shape_ok(DIM, NP, BLEND, 3)) launch_rhs_fb<DIM, NP, BLEND, 256>(c, a, s);
        break;
    default:
        launch_rhs_fb<DIM, NP, BLEND, 0>(c, a, s);
        break;
    }
    return hip_check(c, hipGetLastError(), "k_rhs launch");
}

KArgs make_kargs(const prandtl_ctx *c, const double *y, const double *u0, double *out, int mode,
                 double dt, double c0, double c1, const double *dtp)
{
    KArgs a;
    a.tab = c->tab; a.geo = c->geo; a.phys = c->phys;
    a.blend = c->prm.blending; a.parabolic = c->prm.parabolic;
    a.ec = c->prm.ec_surface; a.quirks = c->prm.quirks;
    a.mode = mode; a.dt = dt; a.c0 = c0; a.c1 = c1;
    a.dtp = dtp;
    a.y = y; a.u0 = u0; a.out = out;
    a.visc = c->d_visc; a.fvf = c->d_fvf; a.alpha = c->d_alpha; a.eta = c->d_eta;
    a.lam = c->d_lam;
    a.skip = c->skip;
    return a;
}

template <int DIM, int NP>
int launch_rhs(prandtl_ctx *c, const double *y, const double *u0, double *out, int mode, double dt,
               double c0, double c1, hipStream_t s)
{
    const KArgs a = make_kargs(c, y, u0, out, mode, dt, c0, c1, c->dtp);
    if (c->prm.blending) return launch_rhs_t<DIM, NP, true>(c, a, s);
    return launch_rhs_t<DIM, NP, false>(c, a, s);
}

template <int DIM, int NP>
int stage_t(prandtl_ctx *c, const double *y, const double *u0, double *out, int mode, double dt,
            double c0, double c1, hipStream_t s)
{
    if (c->prm.parabolic) {
        int rc = launch_visc<DIM, NP>(c, y, nullptr, s);
        if (rc) return rc;
    }
    return launch_rhs<DIM, NP>(c, y, u0, out, mode, dt, c0, c1, s);
}

#define PRANDTL_DISPATCH(FN, ...)                                                   \
    switch (c->dim * 16 + c->Np) {                                                  \
    case 1 * 16 + 2: return FN<1, 2>(__VA_ARGS__);                                  \
    case 1 * 16 + 3: return FN<1, 3>(__VA_ARGS__);                                  \
    case 1 * 16 + 4: return FN<1, 4>(__VA_ARGS__);                                  \
    case 1 * 16 + 5: return FN<1, 5>(__VA_ARGS__);                                  \
    case 1 * 16 + 6: return FN<1, 6>(__VA_ARGS__);                                  \
    case 1 * 16 + 7: return FN<1, 7>(__VA_ARGS__);                                  \
    case 1 * 16 + 8: return FN<1, 8>(__VA_ARGS__);                                  \
    case 2 * 16 + 2: return FN<2, 2>(__VA_ARGS__);                                  \
    case 2 * 16 + 3: return FN<2, 3>(__VA_ARGS__);                                  \
    case 2 * 16 + 4: return FN<2, 4>(__VA_ARGS__);                                  \
    case 2 * 16 + 5: return FN<2, 5>(__VA_ARGS__);                                  \
    case 2 * 16 + 6: return FN<2, 6>(__VA_ARGS__);                                  \
    case 3 * 16 + 3: return FN<3, 3>(__VA_ARGS__);                                  \
    case 3 * 16 + 4: return FN<3, 4>(__VA_ARGS__);                                  \
    case 3 * 16 + 5: return FN<3, 5>(__VA_ARGS__);                                  \
    case 3 * 16 + 6: return FN<3, 6>(__VA_ARGS__);                                  \
    case 3 * 16 + 7: return FN<3, 7>(__VA_ARGS__);                                  \
    case 3 * 16 + 8: return FN<3, 8>(__VA_ARGS__);                                  \
    default: return fail(c, PRANDTL_ERR_UNSUPPORTED, "dim/order not built");        \
    }

int stage(prandtl_ctx *c, const double *y, const double *u0, double *out, int mode, double dt,
          double c0, double c1, hipStream_t s)
{
    PRANDTL_DISPATCH(stage_t, c, y, u0, out, mode, dt, c0, c1, s)
}

int visc_only(prandtl_ctx *c, const double *y, double *grad, hipStream_t s)
{
    PRANDTL_DISPATCH(launch_visc, c, y, grad, s)
}

// k_rhs alone (mode 0, RHS into out); k_visc's outputs must be current
int rhs_only(prandtl_ctx *c, const double *y, double *out, hipStream_t s)
{
    PRANDTL_DISPATCH(launch_rhs, c, y, nullptr, out, 0, 0.0, 0.0, 0.0, s)
}

bool shape_ok(const prandtl_ctx *c, int kernel, int shape)
{
    if (kernel == PRANDTL_KERNEL_VISC)
        return c->prm.parabolic && (shape == 0 || (!c->geo.bcmask && !c->gen && visc_shape_ok(c->dim, c->Np, shape)));
    if (kernel != PRANDTL_KERNEL_RHS) return false;
    if (shape == 0) return true;
    if (c->dim == 3 && c->prm.parabolic && c->prm.quirks) return false;
    if (c->geo.bcmask || c->gen) return false;
    return rhs_shape_ok(c->dim, c->Np, c->prm.blending != 0, shape);
}

void drop_graph(prandtl_ctx *c)
{
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    c->graph_u = nullptr;
}

} // namespace

extern "C" {

int prandtl_lgl(int order, double *x, double *w)
{
    if (order < 1 || order > 16 || !x || !w) return PRANDTL_ERR_ARG;
    std::vector<double> xv, wv;
    host::lgl(order, xv, wv);
    std::memcpy(x, xv.data(), xv.size() * sizeof(double));
    std::memcpy(w, wv.data(), wv.size() * sizeof(double));
    return PRANDTL_OK;
}

} // extern "C"

namespace {

// prandtl_create, prandtl_create_quad and prandtl_create_hex: qm and hm
// null for a Cartesian box (mesh), else a quadrilateral or hexahedral mesh
// (mesh then only carries dim and the order).
int create_impl(const prandtl_mesh_desc *mesh, const prandtl_quad_mesh *qm, const prandtl_params *prm,
                prandtl_ctx **out, const prandtl_hex_mesh *hm = nullptr)
{
    if (!mesh || !prm || !out) return fail(nullptr, PRANDTL_ERR_ARG, "null argument");
    *out = nullptr;
    const int dim = mesh->dim, N = mesh->order;
    if (dim < 1 || dim > 3) return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "dim must be 1, 2 or 3");
    if (N < (dim == 3 ? 2 : 1) || N > (dim == 2 ? 5 : 7))
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "order not built for this dim");
    if (dim == 1 && prm->parabolic)
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "1-D is built for the Euler equations");
    if (prm->precision != PRANDTL_FP64 && prm->precision != PRANDTL_MIXED)
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "precision must be PRANDTL_FP64 or PRANDTL_MIXED");
    if (prm->precision == PRANDTL_MIXED && prm->parabolic && prm->quirks)
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "the mixed form is built for corrected mode (quirks = 0)");
    if (!(prm->gamma > 1.0)) return fail(nullptr, PRANDTL_ERR_ARG, "gamma must exceed 1");
    if (prm->parabolic && !(prm->Pr > 0.0)) return fail(nullptr, PRANDTL_ERR_ARG, "Pr must be positive");
    if (prm->parabolic && prm->sutherland && !(prm->R_gas > 0.0))
        return fail(nullptr, PRANDTL_ERR_ARG, "R_gas must be positive with Sutherland viscosity");
    long long NE = 1;
    if (qm) {
        if (qm->nelem < 1) return fail(nullptr, PRANDTL_ERR_ARG, "empty mesh");
        NE = qm->nelem;
    }
    if (hm) {
        if (hm->nelem < 1) return fail(nullptr, PRANDTL_ERR_ARG, "empty mesh");
        if (N > 4) return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "hexahedral meshes are built for orders 2..4");
        NE = hm->nelem;
    }
    const bool umesh = qm || hm;   // an unstructured mesh, not a box
    for (int r = 0; r < dim && !umesh; r++) {
        if (mesh->nelem[r] < 2) return fail(nullptr, PRANDTL_ERR_ARG, "need >= 2 elements per direction");
        if (!(mesh->hi[r] > mesh->lo[r])) return fail(nullptr, PRANDTL_ERR_ARG, "empty box");
        NE *= mesh->nelem[r];
    }
    if (NE > (1LL << 30)) return fail(nullptr, PRANDTL_ERR_ARG, "mesh too large");
    for (int r = 0; r < dim && !umesh; r++) {
        for (int side = 0; side < 2; side++) {
            const int t = mesh->bc[r][side];
            if (t < PRANDTL_BC_PERIODIC || t > PRANDTL_BC_NO_SLIP_ADIABATIC)
                return fail(nullptr, PRANDTL_ERR_ARG, "unknown boundary type");
            if (prm->parabolic && (t == PRANDTL_BC_SLIP || t == PRANDTL_BC_SYMMETRY))
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED,
                            "slip and symmetry sides have no viscous flux (Navier-Stokes: no-slip walls)");
            if (!prm->parabolic && t == PRANDTL_BC_NO_SLIP_ADIABATIC)
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "no-slip walls need parabolic = 1");
            if (prm->quirks && t == PRANDTL_BC_NO_SLIP_ADIABATIC && mesh->bc_state[r][side][3] != 0.0)
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED,
                            "faithful mode reproduces no-slip walls with zero heat flux only");
            if (t == PRANDTL_BC_SUPERSONIC_INFLOW || t == PRANDTL_BC_SPECIFIED_STATE) {
                const double *g = mesh->bc_state[r][side];
                double ke = 0.0;
                for (int d = 0; d < dim; d++) ke += g[1 + d] * g[1 + d];
                const double pg = (prm->gamma - 1.0) * (g[dim + 1] - 0.5 * ke / g[0]);
                if (!(g[0] > 0.0) || !(pg > 0.0) || !std::isfinite(pg))
                    return fail(nullptr, PRANDTL_ERR_ARG, "boundary state needs positive density and pressure");
            }
        }
        if ((mesh->bc[r][0] == PRANDTL_BC_PERIODIC) != (mesh->bc[r][1] == PRANDTL_BC_PERIODIC))
            return fail(nullptr, PRANDTL_ERR_ARG, "a direction is periodic on both sides or on neither");
    }
    if (prm->axisymmetric && !qm)
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "axisymmetric mode is built for quadrilateral meshes");
    if (prm->axisymmetric && prm->parabolic)
        return fail(nullptr, PRANDTL_ERR_UNSUPPORTED,
                    "axisymmetric mode is built for Euler (the reference weights only the inviscid terms by r)");
    std::vector<double> qstates;   // unstructured meshes: outer states per attribute
    if (umesh) {
        const int nattr = qm ? qm->nattr : hm->nattr;
        const int *attr_bc = qm ? qm->attr_bc : hm->attr_bc;
        const double *attr_state = qm ? qm->attr_state : hm->attr_state;
        const int NS = qm ? 4 : 5;   // values per attribute state
        if (nattr < 0 || (nattr > 0 && !attr_bc)) return fail(nullptr, PRANDTL_ERR_ARG, "bad attribute table");
        qstates.assign(5 * (size_t)std::max(nattr, 1), 0.0);
        for (int a = 0; a < nattr; a++) {
            const int t = attr_bc[a];
            if (t < PRANDTL_BC_PERIODIC || t > PRANDTL_BC_AXIS)
                return fail(nullptr, PRANDTL_ERR_ARG, "unknown boundary type");
            if (prm->parabolic && (t == PRANDTL_BC_SLIP || t == PRANDTL_BC_SYMMETRY))
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED,
                            "slip and symmetry sides have no viscous flux (Navier-Stokes: no-slip walls)");
            if (!prm->parabolic && t == PRANDTL_BC_NO_SLIP_ADIABATIC)
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "no-slip walls need parabolic = 1");
            if (hm && t == PRANDTL_BC_AXIS)
                return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "axis boundaries are for 2-D axisymmetric meshes");
            if (t == PRANDTL_BC_NO_SLIP_ADIABATIC) {
                if (!attr_state) return fail(nullptr, PRANDTL_ERR_ARG, "wall state missing");
                const double *g = attr_state + NS * (size_t)a;
                if (prm->quirks && g[3] != 0.0)
                    return fail(nullptr, PRANDTL_ERR_UNSUPPORTED,
                                "faithful mode reproduces no-slip walls with zero heat flux only");
                for (int v = 0; v < NS; v++) qstates[5 * (size_t)a + v] = g[v];
            }
            if (t == PRANDTL_BC_SUPERSONIC_INFLOW || t == PRANDTL_BC_SPECIFIED_STATE) {
                if (!attr_state) return fail(nullptr, PRANDTL_ERR_ARG, "boundary state missing");
                const double *g = attr_state + NS * (size_t)a;
                double ke = 0.0;
                for (int d = 0; d < dim; d++) ke += g[1 + d] * g[1 + d];
                const double pg = (prm->gamma - 1.0) * (g[dim + 1] - 0.5 * ke / g[0]);
                if (!(g[0] > 0.0) || !(pg > 0.0) || !std::isfinite(pg))
                    return fail(nullptr, PRANDTL_ERR_ARG, "boundary state needs positive density and pressure");
                for (int v = 0; v < NS; v++) qstates[5 * (size_t)a + v] = g[v];
            }
        }
    }
    prandtl_ctx *c = new (std::nothrow) prandtl_ctx();
    if (!c) return fail(nullptr, PRANDTL_ERR_NOMEM, "host allocation");
    c->dim = dim; c->order = N; c->Np = N + 1; c->nv = dim + 2;
    c->mesh = *mesh;
    c->K = dim == 1 ? c->Np : dim == 2 ? c->Np * c->Np : c->Np * c->Np * c->Np;
    c->Nfp = dim == 1 ? 1 : dim == 2 ? c->Np : c->Np * c->Np;
    c->NE = (int)NE;
    c->prm = *prm;
    const int Np = c->Np;

    // 1-D operators (DGSEMIntegrator.cpp:45-80)
    std::vector<double> x, w;
    host::lgl(N, x, w);
    std::vector<double> wb(Np, 1.0), D(Np * Np, 0.0);
    for (int i = 1; i < Np; i++)
        for (int j = 0; j < i; j++) {
            wb[j] *= (x[j] - x[i]);
            wb[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i < Np; i++) wb[i] = 1.0 / wb[i];
    for (int iL = 0; iL < Np; iL++)
        for (int i = 0; i < Np; i++)
            if (iL != i) {
                D[i * Np + iL] = wb[iL] / wb[i] / (x[i] - x[iL]);
                D[i * Np + i] -= D[i * Np + iL];
            }
    std::memset(&c->tab, 0, sizeof(Tables));
    for (int i = 0; i < Np; i++)
        for (int j = 0; j < Np; j++) {
            c->tab.D[i * MAXNP + j] = D[i * Np + j];
            c->tab.Dh[i * MAXNP + j] = D[i * Np + j];
            c->tab.Dh2[i * MAXNP + j] = D[i * Np + j] * 2.0;
        }
    c->tab.Dh[0] += 1.0 / w[0];
    c->tab.Dh[N * MAXNP + N] -= 1.0 / w[N];
    c->tab.Dh2[0] += 1.0 / w[0];
    c->tab.Dh2[N * MAXNP + N] -= 1.0 / w[N];
    // modal basis: 1-D shifted Legendre Vandermonde (ModalBasis.cpp:110-135)
    std::vector<double> V(Np * Np), Vi;
    for (int i = 0; i < Np; i++) {
        std::vector<double> L(Np);
        L[0] = 1.0;
        const double z = 2.0 * x[i] - 1.0;
        if (N >= 1) L[1] = z;
        for (int k = 1; k < N; k++) L[k + 1] = ((2 * k + 1) * z * L[k] - k * L[k - 1]) / (k + 1);
        for (int a = 0; a < Np; a++) V[i * Np + a] = L[a];
    }
    if (!host::invert(Np, V, Vi)) {
        delete c;
        return fail(nullptr, PRANDTL_ERR_ARG, "singular Vandermonde");
    }
    for (int a = 0; a < Np; a++)
        for (int i = 0; i < Np; i++) c->tab.V1i[a * MAXNP + i] = Vi[a * Np + i];

    for (int i = 0; i < Np; i++) {
        c->tab.w[i] = w[i];
        c->tab.x[i] = x[i];
    }
    quad::Geometry qg;
    if (qm) {
        std::string err;
        if (!quad::build(qm, N, x, w, D, prm->axisymmetric != 0, qg, err)) {
            delete c;
            return fail(nullptr, PRANDTL_ERR_ARG, err);
        }
        c->gen = true;
        c->axi = prm->axisymmetric != 0;
        c->geo.axis_floor = prm->axis_rho_floor > 0.0 ? prm->axis_rho_floor : 1e-12;
        c->hmin = qg.hmin;
    }
    if (hm) {
        std::string err;
        if (!quad::build_hex(hm, N, x, w, D, qg, err)) {
            delete c;
            return fail(nullptr, PRANDTL_ERR_ARG, err);
        }
        // the kernels recompute the per-node metrics from the corners
        qg.met.clear();
        qg.jac.clear();
        c->gen = true;
        c->hmin = qg.hmin;
    }
    // Cartesian geometry: J = det diag(h), adj = diag of cofactors
    double h[3] = {1, 1, 1};
    for (int r = 0; r < dim; r++) {
        h[r] = umesh ? 1.0 : (mesh->hi[r] - mesh->lo[r]) / mesh->nelem[r];
        c->geo.n[r] = umesh ? (r == 0 ? c->NE : 1) : mesh->nelem[r];
    }
    if (!umesh) {
        c->hmin = h[0];
        for (int r = 1; r < dim; r++) c->hmin = h[r] < c->hmin ? h[r] : c->hmin;
    }
    for (int r = dim; r < 3; r++) c->geo.n[r] = 1;
    c->geo.NE = c->NE;
    double J;
    if (dim == 1) {
        // the adjugate of a 1 x 1 Jacobian is 1: unit face normals
        J = h[0];
        c->geo.nrm[0] = 1.0; c->geo.nrm[1] = 0.0; c->geo.nrm[2] = 0.0;
    } else if (dim == 2) {
        J = h[0] * h[1];
        c->geo.nrm[0] = h[1]; c->geo.nrm[1] = h[0]; c->geo.nrm[2] = 0.0;
    } else {
        J = h[0] * (h[1] * h[2]);
        c->geo.nrm[0] = h[1] * h[2]; c->geo.nrm[1] = h[0] * h[2]; c->geo.nrm[2] = h[0] * h[1];
    }
    c->geo.invJ = 1.0 / J;
    c->geo.aface = 1.0 / (w[0] * J);
    c->geo.bcmask = 0;
    c->geo.bcs = nullptr;
    for (int r = 0; r < dim && !umesh; r++)
        for (int side = 0; side < 2; side++) c->geo.bcmask |= mesh->bc[r][side] << (3 * (2 * r + side));
    for (int i = 0; i < Np; i++) c->tab.fvs[i] = 1.0 / (J * w[i]);
    // subcell normals n_{i+1/2} = Ja + sum_{l<=i} w_l sum_q D_lq Ja (lines 1212-1233)
    for (int r = 0; r < dim; r++)
        for (int i = 1; i < Np; i++) {
            double nor = c->geo.nrm[r];
            for (int l = 0; l < i; l++) {
                double tmp = 0.0;
                for (int q = 0; q < Np; q++) tmp += c->geo.nrm[r] * D[l * Np + q];
                tmp *= w[l];
                nor += tmp;
            }
            c->tab.subn[r][i - 1] = nor;
        }

    c->phys.gamma = prm->gamma;
    c->phys.gm1 = prm->gamma - 1.0;
    c->phys.gm1i = 1.0 / c->phys.gm1;
    c->phys.mu = prm->mu;
    c->phys.lam_h = prm->parabolic ? prm->mu * prm->gamma * (1.0 / prm->Pr) : 0.0;
    c->phys.alpha_max = prm->alpha_max;
    c->phys.alpha_min = prm->alpha_min;
    c->phys.thresh = 0.5 * std::pow(10.0, -1.8 * std::pow((double)N, 0.25));
    c->phys.sutherland = prm->parabolic && prm->sutherland;
    c->phys.R_gas = prm->R_gas;
    c->phys.PrInv = 1.0 / prm->Pr;

    const size_t NK = (size_t)c->NE * c->K;
    hipError_t e = hipSuccess;
    if (prm->parabolic) {
        if (e == hipSuccess) e = hipMalloc(&c->d_visc, sizeof(double) * (c->nv - 1) * NK);
        if (e == hipSuccess)
            e = hipMalloc(&c->d_fvf, sizeof(double) * (size_t)c->NE * 2 * dim * (c->nv - 1) * c->Nfp);
    }
    if (e == hipSuccess) e = hipMalloc(&c->d_alpha, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipMalloc(&c->d_eta, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipMalloc(&c->d_ya, sizeof(double) * c->nv * NK);
    if (e == hipSuccess) e = hipMalloc(&c->d_yb, sizeof(double) * c->nv * NK);
    if (e == hipSuccess) e = hipMalloc(&c->d_lam, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipMalloc(&c->d_lam_out, sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->d_flag, sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&c->d_dt, sizeof(double));
    if (e == hipSuccess && c->geo.bcmask) {
        double hb[3][2][5];
        std::memcpy(hb, mesh->bc_state, sizeof(hb));
        e = hipMalloc(&c->d_bcs, sizeof(hb));
        if (e == hipSuccess) e = hipMemcpy(c->d_bcs, hb, sizeof(hb), hipMemcpyHostToDevice);
        c->geo.bcs = c->d_bcs;
    }
    if (e == hipSuccess && umesh) {
        // general geometry tables, one allocation
        auto bytes = [](const auto &v) { return v.size() * sizeof(v[0]); };
        constexpr int NPART = 13;
        const size_t parts[NPART] = {bytes(qg.met), bytes(qg.jac), bytes(qg.fnor), bytes(qg.subn), bytes(qstates),
                                     bytes(qg.fface), bytes(qg.fnb), bytes(qg.fnf), bytes(qg.rnode),
                                     bytes(qg.axm), bytes(qg.axj), bytes(qg.axr), bytes(qg.corner)};
        size_t off[NPART + 1] = {0};
        for (int i = 0; i < NPART; i++) off[i + 1] = off[i] + ((parts[i] + 255) / 256) * 256;
        e = hipMalloc(&c->d_geo, off[NPART]);
        const void *src[NPART] = {qg.met.data(), qg.jac.data(), qg.fnor.data(), qg.subn.data(), qstates.data(),
                                  qg.fface.data(), qg.fnb.data(), qg.fnf.data(), qg.rnode.data(),
                                  qg.axm.data(), qg.axj.data(), qg.axr.data(), qg.corner.data()};
        for (int i = 0; i < NPART && e == hipSuccess; i++)
            if (parts[i])
                e = hipMemcpy(static_cast<char *>(c->d_geo) + off[i], src[i], parts[i], hipMemcpyHostToDevice);
        char *b = static_cast<char *>(c->d_geo);
        c->geo.met = reinterpret_cast<const double *>(b + off[0]);
        c->geo.jac = reinterpret_cast<const double *>(b + off[1]);
        c->geo.fnor = reinterpret_cast<const double *>(b + off[2]);
        c->geo.subn = reinterpret_cast<const double *>(b + off[3]);
        c->geo.bcs = reinterpret_cast<const double *>(b + off[4]);
        c->geo.fface = reinterpret_cast<const int *>(b + off[5]);
        c->geo.fnb = reinterpret_cast<const int *>(b + off[6]);
        c->geo.fnf = reinterpret_cast<const int *>(b + off[7]);
        if (hm) c->geo.corner = reinterpret_cast<const double *>(b + off[12]);
        if (c->axi) {
            c->geo.rnode = reinterpret_cast<const double *>(b + off[8]);
            c->geo.axm = reinterpret_cast<const int *>(b + off[9]);
            c->geo.axj = reinterpret_cast<const int *>(b + off[10]);
            c->geo.axr = reinterpret_cast<const double *>(b + off[11]);
        }
    }
    if (e == hipSuccess && c->axi) e = hipMalloc(&c->d_axU, sizeof(double) * c->nv * NK);
    if (e == hipSuccess) e = hipMemset(c->d_alpha, 0, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipMemset(c->d_eta, 0, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipMemset(c->d_lam, 0, sizeof(double) * c->NE);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_create_err = std::string("device allocation: ") + hipGetErrorString(e);
        prandtl_destroy(c);
        return PRANDTL_ERR_NOMEM;
    }
    *out = c;
    return PRANDTL_OK;
}

} // namespace

extern "C" {

int prandtl_create(const prandtl_mesh_desc *mesh, const prandtl_params *prm, prandtl_ctx **out)
{
    return create_impl(mesh, nullptr, prm, out);
}

int prandtl_create_quad(const prandtl_quad_mesh *mesh, int order, const prandtl_params *params, prandtl_ctx **out)
{
    if (!mesh) return fail(nullptr, PRANDTL_ERR_ARG, "null argument");
    prandtl_mesh_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = 2;
    d.order = order;
    d.nelem[0] = mesh->nelem; d.nelem[1] = 1; d.nelem[2] = 1;
    return create_impl(&d, mesh, params, out);
}

int prandtl_create_hex(const prandtl_hex_mesh *mesh, int order, const prandtl_params *params, prandtl_ctx **out)
{
    if (!mesh) return fail(nullptr, PRANDTL_ERR_ARG, "null argument");
    prandtl_mesh_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = 3;
    d.order = order;
    d.nelem[0] = mesh->nelem; d.nelem[1] = 1; d.nelem[2] = 1;
    return create_impl(&d, nullptr, params, out, mesh);
}

int prandtl_hmin(const prandtl_ctx *c, double *hmin)
{
    if (!c || !hmin) return PRANDTL_ERR_ARG;
    *hmin = c->hmin;
    return PRANDTL_OK;
}

void prandtl_destroy(prandtl_ctx *c)
{
    if (!c) return;
    double *bufs[] = {c->d_visc, c->d_fvf, c->d_alpha, c->d_eta, c->d_ya, c->d_yb, c->d_axU};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (c->d_lam) (void)hipFree(c->d_lam);
    if (c->d_loop) (void)hipFree(c->d_loop);
    if (c->d_lam_out) (void)hipFree(c->d_lam_out);
    if (c->d_flag) (void)hipFree(c->d_flag);
    if (c->d_dt) (void)hipFree(c->d_dt);
    if (c->d_bcs) (void)hipFree(c->d_bcs);
    if (c->d_geo) (void)hipFree(c->d_geo);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    delete c;
}

int prandtl_sizes(const prandtl_ctx *c, int *NE, int *K, int *nv)
{
    if (!c || !NE || !K || !nv) return PRANDTL_ERR_ARG;
    *NE = c->NE; *K = c->K; *nv = c->nv;
    return PRANDTL_OK;
}

int prandtl_rhs(prandtl_ctx *c, const double *u, double *dudt, void *stream)
{
    if (!c || !u || !dudt) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (u == dudt) return fail(c, PRANDTL_ERR_ARG, "u and dudt alias");
    return stage(c, u, nullptr, dudt, 0, 0.0, 0.0, 0.0, (hipStream_t)stream);
}

static int rk3_stages(prandtl_ctx *c, double *u, double dt, hipStream_t s)
{
    // y1 = u + dt L(u); y2 = 3/4 u + 1/4 (y1 + dt L(y1)); u = 1/3 u + 2/3 (y2 + dt L(y2))
    int rc = stage(c, u, nullptr, c->d_ya, 1, dt, 0.0, 0.0, s);
    if (!rc) rc = stage(c, c->d_ya, u, c->d_yb, 2, dt, 3.0 / 4, 1.0 / 4, s);
    if (!rc) rc = stage(c, c->d_yb, u, u, 2, dt, 1.0 / 3, 2.0 / 3, s);
    return rc;
}

int prandtl_rk3_step(prandtl_ctx *c, double *u, double dt, void *stream)
{
    if (!c || !u) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    // The null stream cannot be captured; profiling brackets single launches.
    if (!c->use_graph || c->prof || s == nullptr) return rk3_stages(c, u, dt, s);
    int rc = PRANDTL_OK;
    // dt lives in device memory for the graph; copy it only when it changes
    if (!c->graph_dt_set || c->graph_dt != dt || c->graph_stream != s) {
        hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, s, c->d_dt, dt);
        rc = hip_check(c, hipGetLastError(), "set dt");
        if (rc) return rc;
        c->graph_dt = dt;
        c->graph_dt_set = true;
    }
    if (!c->graph_exec || c->graph_u != u || c->graph_stream != s) {
        if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        rc = hip_check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        if (rc) return rc;
        c->dtp = c->d_dt;
        const int rs = rk3_stages(c, u, dt, s);
        c->dtp = nullptr;
        hipGraph_t g = nullptr;
        rc = hip_check(c, hipStreamEndCapture(s, &g), "end capture");
        if (rs) { if (g) (void)hipGraphDestroy(g); return rs; }
        if (rc) return rc;
        c->graph = g;
        rc = hip_check(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0), "graph instantiate");
        if (rc) return rc;
        c->graph_u = u;
        c->graph_stream = s;
    }
    return hip_check(c, hipGraphLaunch(c->graph_exec, s), "graph launch");
}

int prandtl_set_shape(prandtl_ctx *c, int kernel, int shape)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (!shape_ok(c, kernel, shape)) return fail(c, PRANDTL_ERR_UNSUPPORTED, "launch shape not built for this case");
    int &cur = kernel == PRANDTL_KERNEL_RHS ? c->rhs_shape : c->visc_shape;
    if (shape != cur) drop_graph(c);
    cur = shape;
    return PRANDTL_OK;
}

int prandtl_autotune(prandtl_ctx *c, const double *u, int *shapes_out, double *us_out, void *stream)
{
    if (!c || !u) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = hip_check(c, hipEventCreate(&e0), "event");
    if (!rc) rc = hip_check(c, hipEventCreate(&e1), "event");
    const int prev[2] = {c->rhs_shape, c->visc_shape};
    int best[2] = {prev[0], prev[1]};
    for (int i = 0; us_out && i < 2 * PRANDTL_SHAPES; i++) us_out[i] = -1.0;
    // k_visc first (k_rhs reads its outputs), then k_rhs
    for (int kern = PRANDTL_KERNEL_VISC; kern >= PRANDTL_KERNEL_RHS && !rc; kern--) {
        int &cur = kern == PRANDTL_KERNEL_RHS ? c->rhs_shape : c->visc_shape;
        double best_us = 0.0;
        bool any = false;
        for (int sh = 0; sh < PRANDTL_SHAPES && !rc; sh++) {
            if (!shape_ok(c, kern, sh)) continue;
            cur = sh;
            auto run = [&]() {
                return kern == PRANDTL_KERNEL_RHS ? rhs_only(c, u, c->d_yb, s) : visc_only(c, u, nullptr, s);
            };
            constexpr int warm = 2, reps = 10;
            for (int i = 0; i < warm && !rc; i++) rc = run();
            if (!rc) rc = hip_check(c, hipEventRecord(e0, s), "event record");
            for (int i = 0; i < reps && !rc; i++) rc = run();
            if (!rc) rc = hip_check(c, hipEventRecord(e1, s), "event record");
            if (!rc) rc = hip_check(c, hipEventSynchronize(e1), "event sync");
            float ms = 0.0f;
            if (!rc) rc = hip_check(c, hipEventElapsedTime(&ms, e0, e1), "event time");
            if (rc) break;
            const double us = 1e3 * ms / reps;
            if (us_out) us_out[kern * PRANDTL_SHAPES + sh] = us;
            if (!any || us < best_us) { best[kern] = sh; best_us = us; any = true; }
        }
        // the winner stays in place, so k_rhs is timed after the best k_visc
        if (!rc) cur = best[kern];
        if (!rc && kern == PRANDTL_KERNEL_VISC && c->prm.parabolic) rc = visc_only(c, u, nullptr, s);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        c->rhs_shape = prev[0];
        c->visc_shape = prev[1];
        return rc;
    }
    if (c->rhs_shape != prev[0] || c->visc_shape != prev[1]) drop_graph(c);
    if (shapes_out) { shapes_out[0] = c->rhs_shape; shapes_out[1] = c->visc_shape; }
    return PRANDTL_OK;
}

int prandtl_run(prandtl_ctx *c, double *u, double *t_io, double t_final, double cfl, double dt, int max_steps,
                int check_every, int *steps_out, void *stream)
{
    if (!c || !u || !t_io) return fail(c, PRANDTL_ERR_ARG, "null argument");
    const bool variable = cfl > 0.0;
    if (!variable && !(dt > 0.0)) return fail(c, PRANDTL_ERR_ARG, "need cfl > 0 or dt > 0");
    hipStream_t s = (hipStream_t)stream;
    int rc = PRANDTL_OK;
    if (!c->d_loop) rc = hip_check(c, hipMalloc(&c->d_loop, sizeof(LoopState)), "loop state");
    if (rc) return rc;
    // hmin and (N+1)^2 as the host loop forms them
    const double hmin = c->hmin;
    const double cfl_hmin = cfl * hmin, np1sq = (double)(c->Np * c->Np);
    LoopState L0 = {*t_io, variable ? 0.0 : dt, 0.0, 0, 0, 0};
    if (!variable) L0.dt_real = (dt < t_final - *t_io) ? dt : t_final - *t_io;
    rc = hip_check(c, hipMemcpyAsync(c->d_loop, &L0, sizeof(L0), hipMemcpyHostToDevice, s), "loop init");
    if (!rc && variable) {
        // the initial RHS evaluation that seeds the wave speed (Simulation.cpp:903)
        rc = stage(c, u, nullptr, c->d_yb, 0, 0.0, 0.0, 0.0, s);
        if (!rc) {
            hipLaunchKernelGGL(k_loop_dt, dim3(1), dim3(LOOP_BLOCK), 0, s, c->d_loop, c->d_lam, c->NE, cfl_hmin, np1sq,
                               t_final);
            rc = hip_check(c, hipGetLastError(), "loop dt");
        }
    }
    if (rc) return rc;
    // a batch of steps; steps after the loop is done do nothing
    constexpr int BATCH = 16;
    auto batch = [&]() -> int {
        c->skip = &c->d_loop->done;
        c->dtp = &c->d_loop->dt_real;
        int r2 = PRANDTL_OK;
        for (int k = 0; k < BATCH && !r2; k++) {
            r2 = rk3_stages(c, u, 0.0, s);
            if (!r2) {
                hipLaunchKernelGGL(k_loop_end, dim3(1), dim3(LOOP_BLOCK), 0, s, c->d_loop, c->d_lam, c->NE, cfl_hmin,
                                   np1sq, t_final, variable ? 1 : 0, max_steps);
                r2 = hip_check(c, hipGetLastError(), "loop end");
            }
            if (!r2 && check_every > 0) {
                const size_t NK = (size_t)c->NE * c->K;
                const int nb = (int)((NK + 255) / 256);
                const double *uc = u;
                if (c->axi) {
                    // the recovered state is checked
                    hipLaunchKernelGGL(k_axis, dim3(nb), dim3(256), 0, s, u, c->d_axU, c->geo, c->K, c->skip);
                    uc = c->d_axU;
                }
                if (c->dim == 1)
                    hipLaunchKernelGGL(k_loop_check<1>, dim3(nb), dim3(256), 0, s, c->d_loop, uc, NK, c->phys.gm1, check_every);
                else if (c->dim == 2)
                    hipLaunchKernelGGL(k_loop_check<2>, dim3(nb), dim3(256), 0, s, c->d_loop, uc, NK, c->phys.gm1, check_every);
                else
                    hipLaunchKernelGGL(k_loop_check<3>, dim3(nb), dim3(256), 0, s, c->d_loop, uc, NK, c->phys.gm1, check_every);
                r2 = hip_check(c, hipGetLastError(), "loop check");
            }
        }
        c->skip = nullptr;
        c->dtp = nullptr;
        return r2;
    };
    const bool use_graph = c->use_graph && !c->prof && s != nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (use_graph) {
        rc = hip_check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        if (rc) return rc;
        const int rb = batch();
        rc = hip_check(c, hipStreamEndCapture(s, &g), "end capture");
        if (rb) { if (g) (void)hipGraphDestroy(g); return rb; }
        if (!rc) rc = hip_check(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "graph instantiate");
    }
    LoopState L = L0;
    while (!rc) {
        rc = use_graph ? hip_check(c, hipGraphLaunch(ge, s), "graph launch") : batch();
        if (!rc) rc = hip_check(c, hipMemcpyAsync(&L, c->d_loop, sizeof(L), hipMemcpyDeviceToHost, s), "loop state");
        if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
        if (rc || L.done) break;
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    *t_io = L.t;
    if (steps_out) *steps_out = L.steps;
    if (L.bad)
        return fail(c, PRANDTL_ERR_STATE,
                    "non-finite or non-positive density/pressure after step " + std::to_string(L.steps));
    return PRANDTL_OK;
}

int prandtl_set_graph(prandtl_ctx *c, int enable)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    c->use_graph = enable != 0;
    return PRANDTL_OK;
}

int prandtl_max_wave_speed(prandtl_ctx *c, double *out, int reset, void *stream)
{
    if (!c || !out) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, s, c->d_lam, c->NE, c->d_lam_out);
    double v = 0.0;
    int rc = hip_check(c, hipGetLastError(), "reduce lambda");
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&v, c->d_lam_out, sizeof(v), hipMemcpyDeviceToHost, s), "copy lambda");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    *out = v;
    if (reset) {
        hipLaunchKernelGGL(k_fill, dim3((c->NE + 255) / 256), dim3(256), 0, s, c->d_lam, c->NE, 0.0);
        rc = hip_check(c, hipGetLastError(), "reset lambda");
    }
    return rc;
}

int prandtl_alpha(prandtl_ctx *c, double *alpha, double *eta, void *stream)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    int rc = PRANDTL_OK;
    if (alpha) rc = hip_check(c, hipMemcpyAsync(alpha, c->d_alpha, sizeof(double) * c->NE, hipMemcpyDeviceToHost, s), "copy alpha");
    if (!rc && eta) rc = hip_check(c, hipMemcpyAsync(eta, c->d_eta, sizeof(double) * c->NE, hipMemcpyDeviceToHost, s), "copy eta");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    return rc;
}

int prandtl_check_state(prandtl_ctx *c, const double *u, void *stream)
{
    if (!c || !u) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t NK = (size_t)c->NE * c->K;
    hipLaunchKernelGGL(k_reset_int, dim3(1), dim3(1), 0, s, c->d_flag);
    const int nb = (int)((NK + 255) / 256);
    if (c->axi) {
        // axisymmetric: the recovered state
        hipLaunchKernelGGL(k_axis, dim3(nb), dim3(256), 0, s, u, c->d_axU, c->geo, c->K, nullptr);
        u = c->d_axU;
    }
    if (c->dim == 1) hipLaunchKernelGGL(k_check<1>, dim3(nb), dim3(256), 0, s, u, NK, c->phys.gm1, c->d_flag);
    else if (c->dim == 2) hipLaunchKernelGGL(k_check<2>, dim3(nb), dim3(256), 0, s, u, NK, c->phys.gm1, c->d_flag);
    else hipLaunchKernelGGL(k_check<3>, dim3(nb), dim3(256), 0, s, u, NK, c->phys.gm1, c->d_flag);
    int flag = 0;
    int rc = hip_check(c, hipGetLastError(), "k_check launch");
    if (!rc) rc = hip_check(c, hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, s), "copy flag");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    return flag ? fail(c, PRANDTL_ERR_STATE, "non-finite or non-positive density/pressure") : PRANDTL_OK;
}

int prandtl_axis_recover(prandtl_ctx *c, const double *rU, double *U, void *stream)
{
    if (!c || !rU || !U) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (!c->axi) return fail(c, PRANDTL_ERR_UNSUPPORTED, "not an axisymmetric context");
    const size_t NK = (size_t)c->NE * c->K;
    hipLaunchKernelGGL(k_axis, dim3((unsigned)((NK + 255) / 256)), dim3(256), 0, (hipStream_t)stream, rU, U,
                       c->geo, c->K, nullptr);
    return hip_check(c, hipGetLastError(), "k_axis launch");
}

int prandtl_primitives(prandtl_ctx *c, const double *u, double *prim, void *stream)
{
    if (!c || !u || !prim) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t NK = (size_t)c->NE * c->K;
    const int nb = (int)((NK + 255) / 256);
    if (c->axi) {
        // the output of the recovered state (Simulation.cpp:1041-1044)
        const int rc = prandtl_axis_recover(c, u, c->d_axU, stream);
        if (rc) return rc;
        u = c->d_axU;
    }
    if (c->dim == 1) hipLaunchKernelGGL(k_primitives<1>, dim3(nb), dim3(256), 0, s, u, prim, NK, c->phys.gm1);
    else if (c->dim == 2) hipLaunchKernelGGL(k_primitives<2>, dim3(nb), dim3(256), 0, s, u, prim, NK, c->phys.gm1);
    else hipLaunchKernelGGL(k_primitives<3>, dim3(nb), dim3(256), 0, s, u, prim, NK, c->phys.gm1);
    return hip_check(c, hipGetLastError(), "k_primitives launch");
}

int prandtl_debug_gradients(prandtl_ctx *c, const double *u, double *grad, void *stream)
{
    if (!c || !u || !grad) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (!c->prm.parabolic) return fail(c, PRANDTL_ERR_UNSUPPORTED, "not a parabolic context");
    return visc_only(c, u, grad, (hipStream_t)stream);
}

int prandtl_profile_begin(prandtl_ctx *c)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    c->prof = true;
    c->ev_used = 0;
    c->ev_kind.clear();
    return PRANDTL_OK;
}

int prandtl_profile_end(prandtl_ctx *c, double *ms, int *count)
{
    if (!c || !ms || !count) return fail(c, PRANDTL_ERR_ARG, "null argument");
    ms[0] = ms[1] = 0.0;
    count[0] = count[1] = 0;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        int rc = hip_check(c, hipEventSynchronize(c->ev_pool[k + 1]), "event sync");
        if (rc) return rc;
        float t = 0.0f;
        rc = hip_check(c, hipEventElapsedTime(&t, c->ev_pool[k], c->ev_pool[k + 1]), "elapsed");
        if (rc) return rc;
        const int kind = c->ev_kind[k / 2];
        ms[kind] += t;
        count[kind] += 1;
    }
    c->prof = false;
    c->ev_used = 0;
    return PRANDTL_OK;
}

const char *prandtl_last_error(const prandtl_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

} // extern "C"

// ---------------------------------------------------------------------------
// checkpoints: the layout of Simulation.cpp:1084-1116 (Cycle<n>/ folder,
// checkpoint_cycle_<n>.<rank8>.chk and a JSON with time and cycle), with the
// state as raw FP64 (byNODES) behind a header that must match the mesh and
// a 64-bit FNV-1a checksum at the end, and the time at full precision
// ---------------------------------------------------------------------------
namespace {

struct ChkHeader {
    char magic[8];
    int32_t version;
    int32_t dim, order, nelem[3], nv;
    int64_t count;
};
constexpr char CHK_MAGIC[8] = {'P', 'R', 'D', 'T', 'L', 'C', 'H', 'K'};
constexpr int32_t CHK_VERSION = 1;

uint64_t fnv1a64(const void *data, size_t n)
{
    const unsigned char *p = static_cast<const unsigned char *>(data);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::string chk_dir(const char *dir, int cycle) { return std::string(dir) + "/Cycle" + std::to_string(cycle); }
std::string chk_state(const char *dir, int cycle)
{
    return chk_dir(dir, cycle) + "/checkpoint_cycle_" + std::to_string(cycle) + ".00000000.chk";
}
std::string chk_meta(const char *dir, int cycle)
{
    return chk_dir(dir, cycle) + "/checkpoint_cycle_" + std::to_string(cycle) + ".json";
}

// value of "key": <number> in a small JSON object
bool json_number(const std::string &txt, const char *key, double &out)
{
    const std::string k = std::string("\"") + key + "\"";
    const size_t at = txt.find(k);
    if (at == std::string::npos) return false;
    const size_t colon = txt.find(':', at + k.size());
    if (colon == std::string::npos) return false;
    const char *b = txt.c_str() + colon + 1;
    char *e = nullptr;
    out = std::strtod(b, &e);
    return e != b;
}

int64_t mesh_count(const prandtl_mesh_desc *m, int nv)
{
    int64_t NE = 1, K = 1;
    for (int r = 0; r < m->dim; r++) {
        NE *= m->nelem[r];
        K *= m->order + 1;
    }
    return NE * K * nv;
}

int io_fail(const std::string &msg)
{
    g_create_err = msg;
    return PRANDTL_ERR_IO;
}

} // namespace

extern "C" {

int prandtl_checkpoint_write(const char *dir, int cycle, double t, const prandtl_mesh_desc *mesh, int nv,
                             const double *u)
{
    if (!dir || !mesh || !u || cycle < 0 || nv < 1 || mesh->dim < 1 || mesh->dim > 3)
        return fail(nullptr, PRANDTL_ERR_ARG, "bad arguments");
    const int64_t count = mesh_count(mesh, nv);
    std::error_code ec;
    std::filesystem::create_directories(chk_dir(dir, cycle), ec);
    if (ec) return io_fail("cannot create " + chk_dir(dir, cycle) + ": " + ec.message());
    ChkHeader h{};
    std::memcpy(h.magic, CHK_MAGIC, 8);
    h.version = CHK_VERSION;
    h.dim = mesh->dim;
    h.order = mesh->order;
    for (int r = 0; r < 3; r++) h.nelem[r] = r < mesh->dim ? mesh->nelem[r] : 1;
    h.nv = nv;
    h.count = count;
    const uint64_t sum = fnv1a64(u, sizeof(double) * count);
    const std::string path = chk_state(dir, cycle), tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return io_fail("cannot write " + tmp);
        f.write(reinterpret_cast<const char *>(&h), sizeof(h));
        f.write(reinterpret_cast<const char *>(u), sizeof(double) * count);
        f.write(reinterpret_cast<const char *>(&sum), sizeof(sum));
        if (!f) return io_fail("write failed: " + tmp);
    }
    std::filesystem::rename(tmp, path, ec);
    if (ec) return io_fail("cannot rename " + tmp + ": " + ec.message());
    std::ofstream m(chk_meta(dir, cycle), std::ios::trunc);
    if (!m) return io_fail("cannot write " + chk_meta(dir, cycle));
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", t);
    m << "{\n \"time\": " << buf << ",\n \"cycle\": " << cycle << "\n}\n";
    if (!m) return io_fail("write failed: " + chk_meta(dir, cycle));
    return PRANDTL_OK;
}

int prandtl_checkpoint_read(const char *dir, int cycle, const prandtl_mesh_desc *mesh, int nv, double *u,
                            double *t_out)
{
    if (!dir || !mesh || !u || !t_out || cycle < 0 || nv < 1 || mesh->dim < 1 || mesh->dim > 3)
        return fail(nullptr, PRANDTL_ERR_ARG, "bad arguments");
    std::ifstream m(chk_meta(dir, cycle));
    if (!m) return io_fail("cannot open " + chk_meta(dir, cycle));
    const std::string txt((std::istreambuf_iterator<char>(m)), std::istreambuf_iterator<char>());
    double t = 0.0, cyc = -1.0;
    if (!json_number(txt, "time", t) || !json_number(txt, "cycle", cyc))
        return io_fail("no time or cycle in " + chk_meta(dir, cycle));
    if (cyc != (double)cycle) return io_fail("cycle in " + chk_meta(dir, cycle) + " is not " + std::to_string(cycle));
    const std::string path = chk_state(dir, cycle);
    std::ifstream f(path, std::ios::binary);
    if (!f) return io_fail("cannot open " + path);
    ChkHeader h{};
    if (!f.read(reinterpret_cast<char *>(&h), sizeof(h))) return io_fail("truncated header: " + path);
    if (std::memcmp(h.magic, CHK_MAGIC, 8) != 0 || h.version != CHK_VERSION) return io_fail("not a checkpoint: " + path);
    const int64_t count = mesh_count(mesh, nv);
    bool same = h.dim == mesh->dim && h.order == mesh->order && h.nv == nv && h.count == count;
    for (int r = 0; r < mesh->dim; r++) same = same && h.nelem[r] == mesh->nelem[r];
    if (!same) return io_fail("checkpoint mesh or variables differ from this case: " + path);
    std::vector<double> buf((size_t)count);
    if (!f.read(reinterpret_cast<char *>(buf.data()), sizeof(double) * count)) return io_fail("truncated state: " + path);
    uint64_t sum = 0;
    if (!f.read(reinterpret_cast<char *>(&sum), sizeof(sum))) return io_fail("truncated checksum: " + path);
    if (f.peek() != std::char_traits<char>::eof()) return io_fail("trailing data: " + path);
    if (sum != fnv1a64(buf.data(), sizeof(double) * count)) return io_fail("checksum mismatch: " + path);
    std::memcpy(u, buf.data(), sizeof(double) * count);
    *t_out = t;
    return PRANDTL_OK;
}

int prandtl_checkpoint_save(prandtl_ctx *c, const double *u, const char *dir, int cycle, double t, void *stream)
{
    if (!c || !u || !dir) return fail(c, PRANDTL_ERR_ARG, "null argument");
    std::vector<double> h((size_t)c->nv * c->NE * c->K);
    hipStream_t s = (hipStream_t)stream;
    int rc = hip_check(c, hipMemcpyAsync(h.data(), u, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s), "copy state");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    rc = prandtl_checkpoint_write(dir, cycle, t, &c->mesh, c->nv, h.data());
    if (rc) c->err = g_create_err;
    return rc;
}

int prandtl_checkpoint_load(prandtl_ctx *c, double *u, const char *dir, int cycle, double *t_out, void *stream)
{
    if (!c || !u || !dir || !t_out) return fail(c, PRANDTL_ERR_ARG, "null argument");
    std::vector<double> h((size_t)c->nv * c->NE * c->K);
    double t = 0.0;
    int rc = prandtl_checkpoint_read(dir, cycle, &c->mesh, c->nv, h.data(), &t);
    if (rc) {
        c->err = g_create_err;
        return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    rc = hip_check(c, hipMemcpyAsync(u, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s), "copy state");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (!rc) *t_out = t;
    return rc;
}

} // extern "C"

struct prandtl_ensemble {
    int n = 0;
    std::vector<prandtl_ctx *> ctx;
    std::vector<double *> u;               // caller state pointers the arguments were built for
    std::vector<double> dt;                // step sizes last copied to d_dts
    KArgs *d_args[3] = {nullptr, nullptr, nullptr};
    EnsTile *d_tiles = nullptr;
    int ntiles = 0;
    double *d_dts = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    std::string err;
};

namespace {

int ens_epb(int NP, int blend)
{
    switch (NP * 2 + blend) {
    case 6: return RCfg<2, 3, false, ENS_BLOCK>::EPB;
    case 7: return RCfg<2, 3, true, ENS_BLOCK>::EPB;
    case 8: return RCfg<2, 4, false, ENS_BLOCK>::EPB;
    case 9: return RCfg<2, 4, true, ENS_BLOCK>::EPB;
    case 10: return RCfg<2, 5, false, ENS_BLOCK>::EPB;
    case 11: return RCfg<2, 5, true, ENS_BLOCK>::EPB;
    case 12: return RCfg<2, 6, false, ENS_BLOCK>::EPB;
    case 13: return RCfg<2, 6, true, ENS_BLOCK>::EPB;
    default: return 0;
    }
}

// relative cost of one tile of variant v = 2*(NP-3) + blend: nodes per tile
// times the per-node pair work (NP-1 partners per direction) plus the
// indicator and subcell work of a blended tile
double ens_cost(int v)
{
    const int NP = 3 + v / 2, blend = v & 1;
    const int epb = ens_epb(NP, blend);
    return (double)epb * NP * NP * ((NP - 1) + 4.0 + (blend ? 3.0 : 0.0));
}

int ens_fail(prandtl_ensemble *E, int code, const std::string &msg)
{
    if (E) E->err = msg;
    else g_create_err = msg;
    return code;
}

int ens_hip(prandtl_ensemble *E, hipError_t e, const char *what)
{
    if (e == hipSuccess) return PRANDTL_OK;
    return ens_fail(E, PRANDTL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// per-stage argument arrays for the current state pointers
int ens_bind(prandtl_ensemble *E, double *const *u, hipStream_t s)
{
    std::vector<KArgs> a[3];
    for (int i = 0; i < E->n; i++) {
        prandtl_ctx *c = E->ctx[i];
        const double *dtp = E->d_dts + i;
        a[0].push_back(make_kargs(c, u[i], nullptr, c->d_ya, 1, 0.0, 0.0, 0.0, dtp));
        a[1].push_back(make_kargs(c, c->d_ya, u[i], c->d_yb, 2, 0.0, 3.0 / 4, 1.0 / 4, dtp));
        a[2].push_back(make_kargs(c, c->d_yb, u[i], u[i], 2, 0.0, 1.0 / 3, 2.0 / 3, dtp));
    }
    for (int k = 0; k < 3; k++) {
        int rc = ens_hip(E, hipMemcpyAsync(E->d_args[k], a[k].data(), sizeof(KArgs) * E->n, hipMemcpyHostToDevice, s), "copy args");
        if (rc) return rc;
    }
    int rc = ens_hip(E, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    E->u.assign(u, u + E->n);
    return PRANDTL_OK;
}

int ens_stages(prandtl_ensemble *E, hipStream_t s)
{
    for (int k = 0; k < 3; k++)
        hipLaunchKernelGGL(k_ens, dim3(E->ntiles), dim3(ENS_BLOCK), 0, s, E->d_args[k], E->d_tiles);
    return ens_hip(E, hipGetLastError(), "k_ens launch");
}

} // namespace

extern "C" {

int prandtl_ensemble_create(int ncases, const prandtl_mesh_desc *meshes, const prandtl_params *params,
                            prandtl_ensemble **out)
{
    if (!meshes || !params || !out || ncases < 1) return ens_fail(nullptr, PRANDTL_ERR_ARG, "bad arguments");
    if (ncases > ENS_MAX_CASES) return ens_fail(nullptr, PRANDTL_ERR_ARG, "too many cases");
    *out = nullptr;
    for (int i = 0; i < ncases; i++) {
        if (meshes[i].dim != 2 || meshes[i].order < 2 || meshes[i].order > 5)
            return ens_fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "ensemble cases are 2-D with order 2..5");
        if (params[i].parabolic || params[i].ec_surface)
            return ens_fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "ensemble cases are Euler with the reference surface flux");
        for (int r = 0; r < 2; r++)
            if (meshes[i].bc[r][0] != PRANDTL_BC_PERIODIC || meshes[i].bc[r][1] != PRANDTL_BC_PERIODIC)
                return ens_fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "ensemble cases are periodic boxes");
    }
    prandtl_ensemble *E = new (std::nothrow) prandtl_ensemble();
    if (!E) return ens_fail(nullptr, PRANDTL_ERR_NOMEM, "host allocation");
    E->n = ncases;
    std::vector<EnsTile> tiles;
    for (int i = 0; i < ncases; i++) {
        prandtl_ctx *c = nullptr;
        const int rc = prandtl_create(&meshes[i], &params[i], &c);
        if (rc) {
            prandtl_ensemble_destroy(E);
            return ens_fail(nullptr, rc, "case " + std::to_string(i) + ": " + g_create_err);
        }
        E->ctx.push_back(c);
        const int blend = params[i].blending ? 1 : 0;
        const int epb = ens_epb(c->Np, blend);
        const int groups = (c->NE + epb - 1) / epb;
        for (int g = 0; g < groups; g++) tiles.push_back({i, g, 2 * (c->Np - 3) + blend});
    }
    // most expensive tiles first, so the tail of the dispatch order is made of
    // cheap tiles (longest-processing-time-first list scheduling)
    std::stable_sort(tiles.begin(), tiles.end(),
                     [](const EnsTile &a, const EnsTile &b) { return ens_cost(a.v) > ens_cost(b.v); });
    E->ntiles = (int)tiles.size();
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipMalloc(&E->d_args[k], sizeof(KArgs) * ncases);
    if (e == hipSuccess) e = hipMalloc(&E->d_tiles, sizeof(EnsTile) * E->ntiles);
    if (e == hipSuccess) e = hipMalloc(&E->d_dts, sizeof(double) * ncases);
    if (e == hipSuccess) e = hipMemcpy(E->d_tiles, tiles.data(), sizeof(EnsTile) * E->ntiles, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        g_create_err = std::string("ensemble device allocation: ") + hipGetErrorString(e);
        prandtl_ensemble_destroy(E);
        return PRANDTL_ERR_NOMEM;
    }
    *out = E;
    return PRANDTL_OK;
}

void prandtl_ensemble_destroy(prandtl_ensemble *E)
{
    if (!E) return;
    if (E->exec) (void)hipGraphExecDestroy(E->exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    for (int k = 0; k < 3; k++)
        if (E->d_args[k]) (void)hipFree(E->d_args[k]);
    if (E->d_tiles) (void)hipFree(E->d_tiles);
    if (E->d_dts) (void)hipFree(E->d_dts);
    for (prandtl_ctx *c : E->ctx) prandtl_destroy(c);
    delete E;
}

int prandtl_ensemble_case(prandtl_ensemble *E, int i, prandtl_ctx **ctx)
{
    if (!E || !ctx || i < 0 || i >= E->n) return ens_fail(E, PRANDTL_ERR_ARG, "bad arguments");
    *ctx = E->ctx[i];
    return PRANDTL_OK;
}

int prandtl_ensemble_rk3_step(prandtl_ensemble *E, double *const *u, const double *dt, void *stream)
{
    if (!E || !u || !dt) return ens_fail(E, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    bool rebind = E->u.size() != (size_t)E->n;
    for (int i = 0; !rebind && i < E->n; i++) rebind = E->u[i] != u[i];
    if (rebind) {
        const int rc = ens_bind(E, u, s);
        if (rc) return rc;
        if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
        if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
    }
    // step sizes live in device memory, read by the stage kernels: copy them
    // (by a kernel, so the graph is unaffected) only when they change
    int rc = PRANDTL_OK;
    if (rebind || E->dt.size() != (size_t)E->n || !std::equal(E->dt.begin(), E->dt.end(), dt)) {
        DtArgs da;
        for (int i = 0; i < E->n; i++) da.dt[i] = dt[i];
        hipLaunchKernelGGL(k_set_dts, dim3(1), dim3(ENS_MAX_CASES), 0, s, E->d_dts, E->n, da);
        rc = ens_hip(E, hipGetLastError(), "set dt");
        if (rc) return rc;
        E->dt.assign(dt, dt + E->n);
    }
    if (s == nullptr) return ens_stages(E, s);
    if (!E->exec || E->graph_stream != s) {
        if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
        if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
        rc = ens_hip(E, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        if (rc) return rc;
        const int rs = ens_stages(E, s);
        hipGraph_t g = nullptr;
        rc = ens_hip(E, hipStreamEndCapture(s, &g), "end capture");
        if (rs) { if (g) (void)hipGraphDestroy(g); return rs; }
        if (rc) return rc;
        E->graph = g;
        rc = ens_hip(E, hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0), "graph instantiate");
        if (rc) return rc;
        E->graph_stream = s;
    }
    return ens_hip(E, hipGraphLaunch(E->exec, s), "graph launch");
}

const char *prandtl_ensemble_last_error(const prandtl_ensemble *E) { return E ? E->err.c_str() : g_create_err.c_str(); }

} // extern "C"
