// Device side of the DGSEM RHS (gfx950): tables, launch shapes, the k_rhs
// tile body and the k_rhs / k_visc kernel templates.  Included by the
// library's host translation unit and by the per-order instantiation units.
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
#pragma once
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
    double inv_n0, inv_n01;   // 1 / n[0], 1 / (n[0] n[1]): box element index arithmetic
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
    const double *fxy;    // [NE*4*NFP*2] boundary face points' coordinates (quadrilateral meshes)
    const double *bfn;    // null, or per attribute [16]: a split outer state [kind, n0, n1, n2,
                          // offset, speed, below[5], above[5]] (0 kind: the constant state)
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
    const double *tp;     // if set, the time of the step's first stage (device memory)
    double tc;            // this stage's time is *tp + tc dt (RK3SSPSolver: 0, 1, 1/2)
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

// q = e / d for 0 <= e < 2^31 and d >= 1 given dinv = 1/d: the truncated
// double-precision product is off by at most one, and one remainder test
// corrects it (cheaper than an integer division by a run-time divisor).
__device__ __forceinline__ int quot(int e, int d, double dinv)
{
    int q = (int)((double)e * dinv);
    const int rr = e - q * d;
    q += (rr >= d) ? 1 : 0;
    q -= (rr < 0) ? 1 : 0;
    return q;
}

// (i, j, k) of box element e, elements numbered x fastest
__device__ __forceinline__ void box_index(const Geo &g, int e, int *idx)
{
    const int n01 = g.n[0] * g.n[1];
    const int k = quot(e, n01, g.inv_n01);
    const int rem = e - k * n01;
    const int j = quot(rem, g.n[0], g.inv_n0);
    idx[0] = rem - j * g.n[0];
    idx[1] = j;
    idx[2] = k;
}

// The step size of a stage: the device value when the caller keeps dt on the
// device, else the argument.  Written as a branch on the pointer with the
// loaded value passed through an empty asm: a choice between the two loads
// would be compiled to one flat load through a selected address, and a
// flat load still in flight makes the next memory wait drain every
// outstanding load.
__device__ __forceinline__ double step_dt(const KArgs &A)
{
    double dt = A.dt;
    if (A.dtp) {
        double d = *A.dtp;
        asm volatile("" : "+v"(d));
        dt = d;
    }
    return dt;
}

__device__ __forceinline__ int neighbor(const Geo &g, int e, int r, int side)
{
    int idx[3];
    box_index(g, e, idx);
    idx[r] = side ? (idx[r] + 1 == g.n[r] ? 0 : idx[r] + 1) : (idx[r] == 0 ? g.n[r] - 1 : idx[r] - 1);
    return (idx[2] * g.n[1] + idx[1]) * g.n[0] + idx[0];
}

// a, b or c for r = 0, 1, 2 where r differs between the lanes of a wave:
// the three values pass through an empty asm first, so the choice is between
// values, not a load through a selected address of the argument struct
__device__ __forceinline__ double lane_pick3(double a, double b, double c, int r)
{
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c));
    return r == 0 ? a : r == 1 ? b : c;
}

// neighbor() for callers whose direction r differs between the lanes of a
// wave (one thread per line or per face point of several directions).
// Indexing the argument struct with a per-lane r is a vector load from
// argument memory that the thread waits for before its state loads can
// issue; so is a select between the three fields, which the compiler turns
// into one load through a selected address.  The empty asm takes the three
// scalar-loaded values as they are, and the select is then between values.
__device__ __forceinline__ int neighbor_lane(const Geo &g, int e, int r, int side)
{
    int idx[3];
    box_index(g, e, idx);
    int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    asm volatile("" : "+s"(n0), "+s"(n1), "+s"(n2));
    const int nr = r == 0 ? n0 : r == 1 ? n1 : n2;
    idx[r] = side ? (idx[r] + 1 == nr ? 0 : idx[r] + 1) : (idx[r] == 0 ? nr - 1 : idx[r] - 1);
    return (idx[2] * g.n[1] + idx[1]) * g.n[0] + idx[0];
}

// Boundary type of element e's face (r, side), BC_PERIODIC for a face with
// a neighbour.
__device__ __forceinline__ int boundary_of(const Geo &g, int e, int r, int side)
{
    int ix[3];
    box_index(g, e, ix);
    const int idx = ix[r];
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

// LDS slot of node n in the tile kernel's per-node rows.  A wave's pair jobs
// read the nodes of a line, whose slots are a node stride apart; the LDS
// serves 8-byte reads 32 lanes at a time from 64 four-byte banks, so strides
// that are multiples of 16 doubles put a line's nodes on one or two banks.
// 3-D N = 7 pads one double per x-row and one per xy-plane (strides 1, 9,
// 73).  Other orders keep node n at n: their strides are odd or small, or
// (3-D N = 3, z stride 16: two-way) the index arithmetic of a padded slot
// costs the registers of a wave per SIMD.
template <int DIM, int NP> __host__ __device__ constexpr int row_slot(int n)
{
    if (DIM == 3 && NP == 8) return n + n / 8 + n / 64;
    return n;
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
    // LDS row length of the per-node rows (state, node quantities, adjugate
    // rows): see row_slot
    static constexpr int KU = row_slot<DIM, NP>(K - 1) + 1;
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
    double U[C::EPB][C::NV][C::KU];
    double Q[C::EPB][NQ][C::KU];
    // signed surface contributions (surface phase), then the pair and
    // subcell fluxes of one line chunk (direction phases): one buffer
    // (also the indicator's scratch: 4 K values per element)
    static constexpr int FF = cmx(cmx(C::JF, C::FXN), BLEND ? (4 * C::K + C::NV - 1) / C::NV : 1);
    static_assert(!BLEND || 4 * C::K <= C::NV * FF, "indicator scratch fits the flux buffer");
    double F[C::EPB][C::NV][FF];
    double G[GE][C::NV][GK];
    double Al[C::EPB];
    double Red[(C::BLOCK + 63) / 64];
    double M[GEN ? C::EPB : 1][GEN ? DIM * DIM : 1][GEN ? C::KU : 1];   // per-node adjugate rows (general geometry)
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
                                         RhsSmem<DIM, NP, BLEND, QNS3, FB, GEN, AXI> &S, const int tid = -1)
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
    // tid: the thread's index as the caller sees it (a loop over tiles passes
    // one the compiler cannot hoist out of the loop)
    const int t = tid < 0 ? (int)threadIdx.x : tid;
    // the operator's tables, one entry per thread, loaded together and first
    // (indices clamped, so every lane loads and no load waits for another);
    // they go to LDS once the prologue's loads are in flight
    static_assert(NP * NP <= BLOCK, "one table entry per thread");
    const int tq = t < NP * NP ? t : 0, tf = t < NP ? t : 0, ts = t < DIM * NP ? t : 0;
    const double tDh2 = A.tab.Dh2[(tq / NP) * MAXNP + tq % NP];
    const double tV1i = BLEND ? A.tab.V1i[(tq / NP) * MAXNP + tq % NP] : 0.0;
    const double tFvs = A.tab.fvs[tf];
    const double tSub = A.tab.subn[ts / NP][ts % NP];

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
    const double dt = step_dt(A);
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
            const int en = neighbor_lane(A.geo, eo, r, side);
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

    if (t < NP * NP) {
        sDh2[t] = tDh2;
        if (BLEND) sV1i[t] = tV1i;
    }
    if (t < NP) sFvs[t] = tFvs;
    if (t < DIM * NP) sSub[t / NP][t % NP] = tSub;

    // ---- node phase: state and per-node flux quantities into LDS ----
    double lam = 0.0;
    if (act) {
#pragma unroll
        for (int v = 0; v < NV; v++) sU[el][v][row_slot<DIM, NP>(n)] = ys[v];
        if constexpr (GEN)
#pragma unroll
            for (int c = 0; c < MC; c++) sM[el][c][row_slot<DIM, NP>(n)] = mown[c];
        if constexpr (AXI) sR[el][n] = rown;
        const NodeQ<DIM> q = node_quants<DIM>(ys, P);
#pragma unroll
        for (int d = 0; d < DIM; d++) sQ[el][d][row_slot<DIM, NP>(n)] = q.v[d];
        sQ[el][DIM][row_slot<DIM, NP>(n)] = q.beta;
        if constexpr (BLEND) {
            sQ[el][DIM + 1][row_slot<DIM, NP>(n)] = q.ibeta;
            sQ[el][DIM + 2][row_slot<DIM, NP>(n)] = q.lam;
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
        const double nr = lane_pick3(A.geo.nrm[0], A.geo.nrm[1], A.geo.nrm[2], r);
        if (!GEN && (!PREFETCH || jf != t)) {
            // not prefetched, or more surface points than threads: load now
            const int en = neighbor_lane(A.geo, eo, r, side);
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
        for (int v = 0; v < NV; v++) so[v] = sU[ej][v][row_slot<DIM, NP>(no)];
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
                // the attribute's outer state, copied (a pointer that could
                // aim at either this array or the table would keep the
                // array in scratch memory)
                double bs[5];
#pragma unroll
                for (int v = 0; v < 5; v++) bs[v] = A.geo.bcs[(code >> 4) * 5 + v];
                if (A.geo.bfn && A.geo.bfn[(code >> 4) * 16] != 0.0) {
                    // the attribute's time-dependent outer state at this
                    // point: "below" where n . x - offset - speed t < 0
                    // (SpecifiedStateBdrFaceIntegrator.cpp:24-34), evaluated
                    // left to right without contraction
                    const double *fn = A.geo.bfn + (code >> 4) * 16;
                    const double *xp = A.geo.fxy + (efo * NFP + l) * 2;
                    const double tt = (A.tp ? *A.tp : 0.0) + A.tc * dt;
                    double sd = __dmul_rn(fn[1], xp[0]);
                    sd = __dadd_rn(sd, __dmul_rn(fn[2], xp[1]));
                    sd = __dadd_rn(sd, __dmul_rn(fn[3], 0.0));
                    sd = __dsub_rn(sd, fn[4]);
                    sd = __dsub_rn(sd, __dmul_rn(fn[5], tt));
#pragma unroll
                    for (int v = 0; v < 5; v++) bs[v] = (sd < 0.0) ? fn[6 + v] : fn[11 + v];
                }
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
            // the face's first and second states, element by element (not a
            // pointer to one of the two arrays, which would put both in
            // scratch memory)
            double g1[NV], g2[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                g1[v] = own ? so[v] : pn[v];
                g2[v] = own ? pn[v] : so[v];
            }
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
                // (both values read, then chosen: a choice between the two
                // loads compiles to one flat load through a selected address)
                const double rl = sR[ej][no], rg = A.geo.rnode[gn];
                const double rf = own ? rl : rg;
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
                // the side's outer state as values: a select between the
                // table and the interior state would otherwise become one
                // load through a selected pointer, keeping so in scratch
                double bs[5];
#pragma unroll
                for (int v = 0; v < 5; v++) bs[v] = A.geo.bcs[(2 * r + side) * 5 + v];
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
        double s1[NV], s2[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            s1[v] = side ? so[v] : pn[v];
            s2[v] = side ? pn[v] : so[v];
        }
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

    // the block's wave speed: every contribution (nodes, boundary points) is
    // in by now, and reducing it here keeps it out of the registers the pair
    // phase needs
    block_lam_max<BLOCK>(lam, lam_old, A.lam, sRed, slot);

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
                    qa.rho = sU[ej][0][row_slot<DIM, NP>(na)]; qb.rho = sU[ej][0][row_slot<DIM, NP>(nb)];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { qa.v[d] = sQ[ej][d][row_slot<DIM, NP>(na)]; qb.v[d] = sQ[ej][d][row_slot<DIM, NP>(nb)]; }
                    qa.beta = sQ[ej][DIM][row_slot<DIM, NP>(na)]; qb.beta = sQ[ej][DIM][row_slot<DIM, NP>(nb)];
                    double F[NV];
                    if constexpr (GEN) {
                        // ComputeVolumeFlux with the mean of the two adjugate rows
                        double mm[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) mm[d] = (sM[ej][DIM * r + d][row_slot<DIM, NP>(na)] + sM[ej][DIM * r + d][row_slot<DIM, NP>(nb)]) * 0.5;
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
                    for (int v = 0; v < NV; v++) { s1[v] = sU[ej][v][row_slot<DIM, NP>(na)]; s2[v] = sU[ej][v][row_slot<DIM, NP>(nb)]; }
                    qa.rho = s1[0]; qb.rho = s2[0];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { qa.v[d] = sQ[ej][d][row_slot<DIM, NP>(na)]; qb.v[d] = sQ[ej][d][row_slot<DIM, NP>(nb)]; }
                    qa.beta = sQ[ej][DIM][row_slot<DIM, NP>(na)]; qb.beta = sQ[ej][DIM][row_slot<DIM, NP>(nb)];
                    qa.ibeta = sQ[ej][NQ - 2][row_slot<DIM, NP>(na)]; qb.ibeta = sQ[ej][NQ - 2][row_slot<DIM, NP>(nb)];
                    qa.lam = sQ[ej][NQ - 1][row_slot<DIM, NP>(na)]; qb.lam = sQ[ej][NQ - 1][row_slot<DIM, NP>(nb)];
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
            const double ke = 0.5 * ((sU[el][1][row_slot<DIM, NP>(n)] * sU[el][1][row_slot<DIM, NP>(n)] + sU[el][2][row_slot<DIM, NP>(n)] * sU[el][2][row_slot<DIM, NP>(n)]) / sU[el][0][row_slot<DIM, NP>(n)]);
            src = P.gm1 * (sU[el][3][row_slot<DIM, NP>(n)] - ke);
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
            const double yv = AXI ? yw[v] : sU[el][v][row_slot<DIM, NP>(n)];
            double o;
            if (A.mode == 0) o = tot;
            else if (A.mode == 1) o = yv + dt * tot;
            else o = A.c0 * u0v[v] + A.c1 * (yv + dt * tot);
            A.out[gi] = o;
        }
    }
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
            int en = neighbor_lane(A.geo, eo, r, side);
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
    // every start load has been used by now; saying so (free here, before
    // any store) keeps the compiler's count from treating a load left unused
    // on some lane's path as pending, which would make it wait for this
    // thread's face-trace stores to complete further down.  (gfx9's s_waitcnt
    // immediate: vmcnt 0, expcnt 7, lgkmcnt 15.)
    __builtin_amdgcn_s_waitcnt(0x0F70);

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

// State of a time loop run on the device (Simulation::Run's stepping,
// Simulation.cpp:893-907, 1007-1022): time, step size, the clipped step of
// the current step, done flag and step count.
struct LoopState {
    double t, dt, dt_real;
    int done, steps, bad;
};

} // namespace prandtl
