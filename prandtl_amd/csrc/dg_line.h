// k_rhs_ln: the flux-differencing RHS of periodic Cartesian boxes in 3-D with
// one thread per tensor line (gfx950).
//
// Restates the same operator as k_rhs (DGSEMIntegrator.cpp:151-433,435-494,
// 601-750, 972-1196; PerssonPeraireIndicator.cpp:21-58; DGSEMOperator.cpp:
// 60-80; MFEM's RK3SSPSolver stage arithmetic) with another decomposition:
//
//  - node phase: the workgroup's EPB elements are loaded (coalesced, one
//    node per thread and pass) into LDS with each node's velocity, beta =
//    rho/(2p), 1/beta and wave speed;
//  - line phase: thread (element, direction r, line l) holds the NP nodes of
//    its line in registers, evaluates the line's NP(NP-1)/2 symmetric volume
//    pairs, its NP-1 subcell-FV interfaces (blended elements) and the two
//    surface points at its ends (this element's face points of the faces
//    2r and 2r+1, with the neighbours' traces read from HBM), and sums each
//    node's contributions of direction r in registers;
//  - the DIM partial sums of a node are added in LDS in direction order
//    (DIM barrier-separated passes, no atomics), then the node phase applies
//    1/J, the viscous volume term from k_visc and the RK stage combination.
//
// A face point's flux is computed by both neighbouring elements with the
// face's canonical orientation (elem1 = low side), so both sides see the same
// value, as in k_rhs.  Summation order differs from k_rhs (per direction, then
// over directions), so the two kernels agree to rounding, not bitwise; every
// launch of this kernel rounds identically.
//
// The two-point flux here is ChandrashekarFlux::ComputeVolumeFlux /
// ComputeFaceFlux (ChandrashekarFlux.cpp:11-161) rearranged to two
// reciprocals per pair instead of four divisions:
//   - the log mean is formed as num/den (series branch: 52.5 s t^3 over the
//     series denominator; log branch: (y-x) over log(y/x)), so rho_ln needs
//     one reciprocal of den;
//   - 1/beta_ln = den_b/num_b and p_hat = (rho1+rho2)/(2(beta1+beta2)) share
//     one reciprocal of num_b (beta1+beta2);
//   - the energy flux rho_ln Vn (... + p_hat/rho_ln) is rho_ln Vn (...) +
//     Vn p_hat (the distributive law), so 1/rho_ln is never formed.
// Each reciprocal is v_rcp_f64 refined by two Newton-Raphson steps (the
// textbook iteration r <- r + r(1 - x r)), within an ulp or two of the
// correctly rounded value; the log branch keeps the IEEE quotient y/x inside
// the logarithm.  The results differ from the reference's by rounding only.
#pragma once
#include "dg_device.h"

namespace prandtl {

// Log mean LM(x, y) = num / den, the branches of ComputeLogMean
// (BasicOperations.cpp:6-12, eps 1e-4) as in logmean(): u = d^2/s^2.
__device__ __forceinline__ void lm_parts(double x, double y, double &num, double &den)
{
    const double d = y - x, s = y + x;
    const double q = d * d, t = s * s;
    if (q < 1e-4 * t) {
        num = 52.5 * s * (t * t * t);
        den = lm_series_den(q, t);
    } else {
        num = d;
        den = log(y / x);
    }
}

// Chandrashekar two-point flux for the normal nr e_r (see the file header).
// DISS = false: the volume flux; DISS = true: the face flux with the jumps of
// s1, s2 and dissipation scaled by |nr|; quirk3d: the 3-D energy term of
// ChandrashekarFlux.cpp:131 as written.
template <int DIM, bool DISS>
__device__ __forceinline__ void ec_flux(int r, double nr, const NodeQ<DIM> &a, const NodeQ<DIM> &b,
                                        const double *s1, const double *s2, const Phys &P, bool quirk3d,
                                        double *F)
{
    double nrh, drh, nbt, dbt;
    lm_parts(a.rho, b.rho, nrh, drh);
    lm_parts(a.beta, b.beta, nbt, dbt);
    const double rho_ln = nrh * rcp_nr(drh);
    const double bsum = a.beta + b.beta;
    const double rb = rcp_nr(nbt * bsum);
    const double ibeta_ln = dbt * bsum * rb;
    const double rsum = a.rho + b.rho;
    const double p_hat = 0.5 * rsum * nbt * rb;
    double vm[DIM];
    double h = 0.0, diss = 0.0;
    const double drho = b.rho - a.rho;
    const double rho_mean = 0.5 * rsum;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vm[d] = 0.5 * (a.v[d] + b.v[d]);
        const double hd = -0.25 * (a.v[d] * a.v[d] + b.v[d] * b.v[d]) + vm[d] * vm[d];
        h = (d == 0) ? hd : h + hd;
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
    h += 0.5 * ibeta_ln * P.gm1i;
    const double rVn = rho_ln * Vn;
    F[0] = rVn;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double mom = Vn * (rho_ln * vm[d]);
        F[1 + d] = (d == r) ? mom + p_hat * nr : mom;
    }
    F[DIM + 1] = rVn * h + Vn * p_hat;
    if (DISS) {
        diss += 0.5 * drho * P.gm1i * ibeta_ln + 0.5 * rho_mean * P.gm1i * (b.ibeta - a.ibeta);
        const double lam = (a.lam < b.lam) ? b.lam : a.lam;
        const double hl = 0.5 * lam * fabs(nr);
        F[0] = F[0] - hl * (s2[0] - s1[0]);
#pragma unroll
        for (int d = 0; d < DIM; d++) F[1 + d] = F[1 + d] - hl * (s2[1 + d] - s1[1 + d]);
        F[DIM + 1] = F[DIM + 1] - hl * diss;
    }
}

constexpr int line_epb(int TE)
{
    // elements per workgroup: the best thread utilisation within 256 threads
    int best = 1;
    double bu = 0.0;
    for (int e = 1; r64(e * TE) <= 256; e++) {
        const double u = (double)(e * TE) / r64(e * TE);
        if (u >= bu) { bu = u; best = e; }
    }
    return best;
}

template <int DIM, int NP> struct LCfg {
    static constexpr int NV = DIM + 2;
    static constexpr int K = npow(NP, DIM);
    static constexpr int NFP = npow(NP, DIM - 1);
    static constexpr int NPAIR = NP * (NP - 1) / 2;
    static constexpr int TE = DIM * NFP;          // line threads per element
    static constexpr int EPB = line_epb(TE);
    static constexpr int BLOCK = r64(EPB * TE);
    static constexpr int NQ = DIM + 3;             // v[DIM], beta, 1/beta, lambda
    // SWZ (NP = 4): the skewed LDS slots of line_slot with 16 doubles of
    // padding per element, and half-waves that each hold one direction's
    // lines of two elements; otherwise node n at n + n / NP, and lanes in
    // (element, direction, line) order
    static constexpr bool SWZ = (NP == 4);
    static constexpr int KP = SWZ ? K : K + K / NP;   // doubles per quantity row
    static constexpr int EPAD = SWZ ? 16 : 0;         // doubles after each element's rows
#ifndef PRANDTL_LINE_WAVES
#define PRANDTL_LINE_WAVES 3
#endif
    static constexpr int WAVES = PRANDTL_LINE_WAVES;   // waves per SIMD asked of the register allocator
};

// The orders the line kernel is built for (3-D, N = 2..4).
constexpr bool line_built(int DIM, int NP) { return DIM == 3 && NP >= 3 && NP <= 5; }

// LDS slot of node (i, j, k) in a quantity's row.
//
// NP = 4: slot 16k + 4((j+k) mod 4) + ((i+k) mod 4).  The LDS has 64 banks
// of 4 bytes and serves a wave's 8-byte accesses 32 lanes at a time, so 32
// lanes reading distinct doubles run in one pass when the slots are distinct
// mod 32.  In each of the planes i = m, j = m and k = m (the node m of every
// line of one direction) the slots are distinct mod 16; with element blocks
// 16 (mod 32) doubles apart, a half-wave holding one direction's lines of two
// elements reads 32 distinct banks.  Node-order passes (32 consecutive nodes,
// k in {0, 1} or {2, 3}) are distinct mod 32 as well.
// Other orders: node n = i + NP(j + NP k) at n + n / NP.
template <int NP> __device__ __forceinline__ int line_slot(int i, int j, int k)
{
    if constexpr (NP == 4) {
        return 16 * k + 4 * ((j + k) & 3) + ((i + k) & 3);
    } else {
        const int n = (k * NP + j) * NP + i;
        return n + n / NP;
    }
}
template <int NP> __device__ __forceinline__ int node_slot(int n)
{
    return line_slot<NP>(n % NP, (n / NP) % NP, n / (NP * NP));
}

// QNS3: the reference's NS-3D zeta-partner defect (DGSEMIntegrator.cpp:702):
// the zeta partners m < i of node i take the eta-direction flux of node m
// with its eta-line's last node (0 on that last node) instead of F(i, m).
// VT: element type of k_visc's outputs (float in the mixed form).
template <int DIM, int NP, bool BLEND, bool QNS3, typename VT>
__global__ void __launch_bounds__((LCfg<DIM, NP>::BLOCK))
__attribute__((amdgpu_waves_per_eu((LCfg<DIM, NP>::WAVES)))) k_rhs_ln(const KArgs A)
{
    using C = LCfg<DIM, NP>;
    constexpr int NV = C::NV, K = C::K, NFP = C::NFP, TE = C::TE, EPB = C::EPB, BLOCK = C::BLOCK;
    constexpr int NQ = C::NQ;
    static_assert(DIM == 3, "line kernel: 3-D boxes");
    static_assert(!BLEND || 4 * K <= NV * K, "indicator scratch fits the partial sums");
    // Per element: NV rows of the state (sU) and NQ rows of node quantities
    // (sQ), KP doubles each (node slots of line_slot), then EPAD doubles, so
    // one address per node reaches every row.  The partial sums (sP, written
    // after the line phase) take the rows of sQ unless blending, whose
    // indicator uses sP as scratch while sQ is live.
    constexpr int KP = C::KP;
    constexpr int SE = (NV + NQ) * KP + C::EPAD, SPE = NV * KP + C::EPAD;
    static_assert(NQ >= NV, "partial sums fit the node-quantity rows");
    __shared__ double sUQ[EPB * SE];
    __shared__ double sPb[BLEND ? EPB * SPE : 1];
    auto sU = [&](int e2, int v, int s) -> double & { return sUQ[e2 * SE + v * KP + s]; };
    auto sQ = [&](int e2, int q, int s) -> double & { return sUQ[e2 * SE + (NV + q) * KP + s]; };
    auto sP = [&](int e2, int v, int s) -> double & {
        if constexpr (BLEND) return sPb[e2 * SPE + v * KP + s];
        else return sUQ[e2 * SE + (NV + v) * KP + s];
    };
    __shared__ double sV1i[BLEND ? NP * NP : 1];
    __shared__ double sAl[EPB];
    __shared__ double sRed[(BLOCK + 63) / 64];

    if (A.skip && *A.skip) return;   // uniform: before any barrier
    const int t = threadIdx.x;
    const int ebase = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int NE = A.geo.NE;
    const size_t NK = (size_t)NE * K;
    const Phys &P = A.phys;
    const bool quirk3d = A.quirks != 0;
    const double dt = step_dt(A);
    const double lam_old = (t == 0) ? A.lam[blockIdx.x] : 0.0;
    if (BLEND)
        for (int i = t; i < NP * NP; i += BLOCK) sV1i[i] = A.tab.V1i[(i / NP) * MAXNP + i % NP];

    // ---- this thread's line: element el, direction r, line l ----
    int el, r, l;
    if constexpr (C::SWZ) {
        // half-wave h holds the direction-(h mod 3) lines of the elements
        // 2(h / 3) and 2(h / 3) + 1, 16 lines each
        static_assert(NFP == 16 && EPB % 2 == 0 && BLOCK == EPB * TE, "line kernel: half-wave lane map");
        const int h = t >> 5;
        r = h % 3;
        el = 2 * (h / 3) + ((t >> 4) & 1);
        l = t & 15;
    } else {
        el = t / TE;
        const int rem = t - el * TE;
        r = rem / NFP;
        l = rem - r * NFP;
    }
    const int e = ebase + el;
    const bool lact = (el < EPB) && (e < NE);
    // node of position m on the line: base + m * stride (x fastest); the
    // line's nodes are (m, l mod NP, l / NP), (l mod NP, m, l / NP) or
    // (l mod NP, l / NP, m) for r = 0, 1, 2
    const int base = r == 0 ? l * NP : (r == 1 ? (l / NP) * NP * NP + l % NP : l);
    const int stride = r == 0 ? 1 : (r == 1 ? NP : NP * NP);
    const int la = l % NP, lb = l / NP;
    // unskewed rows: line_slot(base + m stride) = pb + m ps
    const int pb = r == 0 ? l * (NP + 1) : (r == 1 ? lb * NP * (NP + 1) + la : l + lb);
    const int ps = r == 0 ? 1 : (r == 1 ? NP + 1 : NP * (NP + 1));
    // LDS slot of the line's node m: pb + m ps, or with the skewed slots
    // lsl[m], formed at the start of the line phase (formed here, they would
    // be held through the node phase)
    int lsl[NP];
    auto lnode = [&](int m) {
        if constexpr (C::SWZ)
            return lsl[m];
        else
            return pb + m * ps;
    };
    const double nr = r == 0 ? A.geo.nrm[0] : (r == 1 ? A.geo.nrm[1] : A.geo.nrm[2]);

    // Every global load of the start is issued on every path, the workgroup's
    // nodes first.  Loads complete in issue order, so the node phase then
    // waits for its own loads alone while the line's traces are still in
    // flight; a load skipped on some path would make that wait a wait for
    // everything.  Lanes without an element load from index 0 and discard
    // the value; without viscous terms the viscous traces are read from the
    // state array (at most 3 NP^2 entries, within its 5 NP^3) and discarded.
    const int nact = (NE - ebase) < EPB ? (NE - ebase) * K : EPB * K;
    constexpr int NPASS = (EPB * K + BLOCK - 1) / BLOCK;
    double ysp[NPASS][NV];
#pragma unroll
    for (int p = 0; p < NPASS; p++) {
        const int i = t + p * BLOCK;
        const size_t gi = (i < nact) ? (size_t)ebase * K + i : 0;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double y = A.y[v * NK + gi];
            ysp[p][v] = (i < nact) ? y : 1.0;
        }
    }

    // neighbour traces at both ends of the line, in flight during the node
    // phase
    double pl[NV], ph[NV];
    const int enl = lact ? neighbor_lane(A.geo, e, r, 0) : 0, enh = lact ? neighbor_lane(A.geo, e, r, 1) : 0;
    {
        const size_t gl = lact ? (size_t)enl * K + base + (NP - 1) * stride : 0;
        const size_t gh = lact ? (size_t)enh * K + base : 0;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double a = A.y[v * NK + gl], b = A.y[v * NK + gh];
            pl[v] = lact ? a : 1.0;
            ph[v] = lact ? b : 1.0;
        }
    }
    // viscous face traces Fv.n at the line's two face points, this element's
    // and the neighbour's, also in flight during the node phase
    VT fto[2][NV - 1], ftn[2][NV - 1];
    {
        const bool vact = A.parabolic && lact;
        const VT *fv = A.parabolic ? static_cast<const VT *>(A.fvf) : reinterpret_cast<const VT *>(A.y);
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const size_t oo = vact ? (((size_t)e * 2 * DIM + 2 * r + side) * (NV - 1)) * NFP + l : 0;
            const size_t on = vact ? (((size_t)(side ? enh : enl) * 2 * DIM + 2 * r + 1 - side) * (NV - 1)) * NFP + l : 0;
#pragma unroll
            for (int v = 0; v < NV - 1; v++) {
                const VT a = fv[oo + v * NFP], b = fv[on + v * NFP];
                fto[side][v] = vact ? a : VT(0);
                ftn[side][v] = vact ? b : VT(0);
            }
        }
    }

    // ---- node phase: state and per-node flux quantities into LDS ----
    double lam = 0.0;
#pragma unroll
    for (int p = 0; p < NPASS; p++) {
        const int i = t + p * BLOCK;
        if (i >= nact) continue;
        const int e2 = i / K, n = i - e2 * K;
        double ys[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) ys[v] = ysp[p][v];
        const int np = node_slot<NP>(n);
#pragma unroll
        for (int v = 0; v < NV; v++) sU(e2, v, np) = ys[v];
        const NodeQ<DIM> q = node_quants<DIM>(ys, P);
#pragma unroll
        for (int d = 0; d < DIM; d++) sQ(e2, d, np) = q.v[d];
        sQ(e2, DIM, np) = q.beta;
        sQ(e2, DIM + 1, np) = q.ibeta;
        sQ(e2, DIM + 2, np) = q.lam;
        const double lq = (q.lam == q.lam) ? q.lam : 0.0;
        lam = (lam < lq) ? lq : lam;
    }
    __syncthreads();

    // ---- modal indicator and blending coefficient (element-local) ----
    if constexpr (BLEND) {
        double *sX = sPb;
        double *sY = sX + EPB * K;
        for (int i = t; i < nact; i += BLOCK) {
            const int e2 = i / K, n = i - e2 * K;
            double ys[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) ys[v] = sU(e2, v, node_slot<NP>(n));
            // rho * ComputePressure(state) with V /= rho (BasicOperations.cpp:31-37)
            const double rr = 1.0 / ys[0];
            double vv = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                const double v = ys[1 + d] * rr;
                vv = (d == 0) ? v * v : vv + v * v;
            }
            sX[i] = ys[0] * (P.gm1 * (ys[DIM + 1] - 0.5 * ys[0] * vv));
        }
        __syncthreads();
        // modes = V^-1 q, one direction at a time (V^-1 = kron of 1-D inverses)
        double *src = sX, *dst = sY;
#pragma unroll
        for (int rr = 0; rr < DIM; rr++) {
            for (int i = t; i < nact; i += BLOCK) {
                const int e2 = i / K, n = i - e2 * K;
                const int ii[3] = {n % NP, (n / NP) % NP, n / (NP * NP)};
                const int a = ii[rr];
                const int st = rr == 0 ? 1 : (rr == 1 ? NP : NP * NP);
                const int b0 = n - a * st;
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < NP; m++)
                    s = (m == 0) ? sV1i[a * NP + m] * src[e2 * K + b0 + m * st]
                                 : s + sV1i[a * NP + m] * src[e2 * K + b0 + m * st];
                dst[i] = s;
            }
            __syncthreads();
            double *tmp = src; src = dst; dst = tmp;
        }
        // modal energies: all modes, degrees <= N-1 and <= N-2 in every
        // direction (PerssonPeraireIndicator.cpp:38-52), summed per element
        // by the same binary tree as k_rhs
        double *sE0 = dst, *sE1 = sX + 2 * EPB * K, *sE2 = sX + 3 * EPB * K;
        for (int i = t; i < nact; i += BLOCK) {
            const int n = i % K;
            const int ii[3] = {n % NP, (n / NP) % NP, n / (NP * NP)};
            bool in1 = true, in2 = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                if (ii[d] > NP - 3) in2 = false;
                if (ii[d] > NP - 2) in1 = false;
            }
            const double md = src[i];
            const double e2 = md * md;
            sE0[i] = e2;
            sE1[i] = in1 ? e2 : 0.0;
            sE2[i] = in2 ? e2 : 0.0;
        }
        __syncthreads();
        constexpr int HALF = pow2_at_least(K) / 2;
#pragma unroll
        for (int hh = HALF; hh >= 1; hh >>= 1) {
            for (int i = t; i < nact; i += BLOCK) {
                const int n = i % K;
                if (n < hh && n + hh < K) {
                    sE0[i] += sE0[i + hh];
                    sE1[i] += sE1[i + hh];
                    sE2[i] += sE2[i + hh];
                }
            }
            __syncthreads();
        }
        for (int e2 = t; e2 < EPB; e2 += BLOCK) {
            if (ebase + e2 >= NE) continue;
            const double mm = sE0[e2 * K], m1 = sE1[e2 * K], m2 = sE2[e2 * K];
            double eta = 1.0 - m1 / mm;
            const double eta2 = 1.0 - m2 / m1;
            eta = (eta < eta2) ? eta2 : eta;
            // DGSEMOperator.cpp:67-77, sharpness 9.21024 (DGSEMOperator.hpp:50)
            double al = 1.0 / (1.0 + exp(-9.21024 * (eta - P.thresh) / P.thresh));
            if (al < P.alpha_min) al = 0.0;
            else if (al > 1.0 - P.alpha_min) al = 1.0;
            al = (P.alpha_max < al) ? P.alpha_max : al;
            sAl[e2] = al;
            A.alpha[ebase + e2] = al;
            A.eta[ebase + e2] = eta;
        }
        __syncthreads();
    }

    // ---- line phase ----
    // acc[a] collects node a's direction-r terms already scaled: surface
    // terms, (-(1 - alpha) / J) 2 D_ab F for the volume pairs, alpha (G_{a-1/2}
    // - G_{a+1/2}) / (J w_a) for the subcell interfaces
    double acc[NP][NV];
#pragma unroll
    for (int a = 0; a < NP; a++)
#pragma unroll
        for (int v = 0; v < NV; v++) acc[a][v] = 0.0;
    if constexpr (C::SWZ) {
#pragma unroll
        for (int m = 0; m < NP; m++)
            lsl[m] = line_slot<NP>(r == 0 ? m : la, r == 1 ? m : (r == 0 ? la : lb), r == 2 ? m : lb);
    }
    if (lact) {
        const double aface = A.geo.aface;
        // surface points: low face (neighbour = elem1, normal +e_r) and high
        // face (this element = elem1)
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int no = lnode(side ? NP - 1 : 0);
            double so[NV], pn[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) { so[v] = sU(el, v, no); pn[v] = side ? ph[v] : pl[v]; }
            double *s1 = side ? so : pn;
            double *s2 = side ? pn : so;
            double fvn[NV];
            fvn[0] = 0.0;
            bool own_q = true;
            if (A.parabolic) {
                // viscous face traces Fv.n of both sides of this face point
#pragma unroll
                for (int v = 1; v < NV; v++)
                    fvn[v] = (((double)fto[side][v - 1] + (double)ftn[side][v - 1]) * 0.5) * nr;
                if (A.quirks) {
                    // ComputeViscousFlux left velocities in the momentum slots
                    // of state1/state2 (DGSEMIntegrator.cpp:471-481)
                    const double r1 = 1.0 / s1[0], r2 = 1.0 / s2[0];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { s1[1 + d] *= r1; s2[1 + d] *= r2; }
                    own_q = false;
                }
            }
            NodeQ<DIM> qo;
            if (own_q) {
                qo.rho = so[0];
#pragma unroll
                for (int d = 0; d < DIM; d++) qo.v[d] = sQ(el, d, no);
                qo.beta = sQ(el, DIM, no);
                qo.ibeta = sQ(el, DIM + 1, no);
                qo.lam = sQ(el, DIM + 2, no);
            } else {
                qo = node_quants<DIM>(so, P);
            }
            const NodeQ<DIM> qn = node_quants<DIM>(pn, P);
            double F[NV];
            if (side) {
                if (A.ec) ec_flux<DIM, false>(r, nr, qo, qn, nullptr, nullptr, P, false, F);
                else ec_flux<DIM, true>(r, nr, qo, qn, s1, s2, P, quirk3d, F);
#pragma unroll
                for (int v = 0; v < NV; v++) acc[NP - 1][v] = aface * (A.parabolic ? (-F[v] + fvn[v]) : -F[v]);
            } else {
                if (A.ec) ec_flux<DIM, false>(r, nr, qn, qo, nullptr, nullptr, P, false, F);
                else ec_flux<DIM, true>(r, nr, qn, qo, s1, s2, P, quirk3d, F);
#pragma unroll
                for (int v = 0; v < NV; v++) acc[0][v] = aface * (A.parabolic ? (F[v] - fvn[v]) : F[v]);
            }
        }

        const double alpha_e = BLEND ? sAl[el] : 0.0;
        const double cv = BLEND ? -(1.0 - alpha_e) * A.geo.invJ : -A.geo.invJ;
        const bool zq = QNS3 && r == 2;
        // symmetric volume pairs a < b: 2 D_ab F to node a, 2 D_ba F to node b
#pragma unroll
        for (int a = 0; a < NP; a++) {
            NodeQ<DIM> qa;
            const int na = lnode(a);
            qa.rho = sU(el, 0, na);
#pragma unroll
            for (int d = 0; d < DIM; d++) qa.v[d] = sQ(el, d, na);
            qa.beta = sQ(el, DIM, na);
#pragma unroll
            for (int b = a + 1; b < NP; b++) {
                NodeQ<DIM> qb;
                const int nb = lnode(b);
                qb.rho = sU(el, 0, nb);
#pragma unroll
                for (int d = 0; d < DIM; d++) qb.v[d] = sQ(el, d, nb);
                qb.beta = sQ(el, DIM, nb);
                double F[NV];
                ec_flux<DIM, false>(r, nr, qa, qb, nullptr, nullptr, P, false, F);
                const double dab = cv * A.tab.Dh2[a * MAXNP + b], dba = cv * A.tab.Dh2[b * MAXNP + a];
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    acc[a][v] += dab * F[v];
                    if (!zq) acc[b][v] += dba * F[v];
                }
            }
            if (QNS3 && zq && a < NP - 1) {
                // zeta partners a < b of node b: the eta flux g of node a with
                // node (i, NP-1, a) of its eta line; 0 on the eta line's last node
                const int i0 = l % NP, j0 = l / NP;
                if (j0 < NP - 1) {
                    const int ne = line_slot<NP>(i0, NP - 1, a);
                    NodeQ<DIM> qe;
                    qe.rho = sU(el, 0, ne);
#pragma unroll
                    for (int d = 0; d < DIM; d++) qe.v[d] = sQ(el, d, ne);
                    qe.beta = sQ(el, DIM, ne);
                    double g[NV];
                    ec_flux<DIM, false>(1, A.geo.nrm[1], qa, qe, nullptr, nullptr, P, false, g);
#pragma unroll
                    for (int b = a + 1; b < NP; b++) {
                        const double dba = cv * A.tab.Dh2[b * MAXNP + a];
#pragma unroll
                        for (int v = 0; v < NV; v++) acc[b][v] += dba * g[v];
                    }
                }
            }
        }
        if constexpr (BLEND) {
            if (alpha_e > 0.0) {
                // subcell FV along the line: G at the interfaces between
                // consecutive nodes, 0 at the element boundary
                // (DGSEMIntegrator.cpp:972-1196)
                double gprev[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) gprev[v] = 0.0;
#pragma unroll
                for (int i = 0; i < NP - 1; i++) {
                    const int na = lnode(i), nb = lnode(i + 1);
                    double s1[NV], s2[NV];
#pragma unroll
                    for (int v = 0; v < NV; v++) { s1[v] = sU(el, v, na); s2[v] = sU(el, v, nb); }
                    NodeQ<DIM> qa, qb;
                    qa.rho = s1[0]; qb.rho = s2[0];
#pragma unroll
                    for (int d = 0; d < DIM; d++) { qa.v[d] = sQ(el, d, na); qb.v[d] = sQ(el, d, nb); }
                    qa.beta = sQ(el, DIM, na); qb.beta = sQ(el, DIM, nb);
                    qa.ibeta = sQ(el, DIM + 1, na); qb.ibeta = sQ(el, DIM + 1, nb);
                    qa.lam = sQ(el, DIM + 2, na); qb.lam = sQ(el, DIM + 2, nb);
                    const double si = r == 0 ? A.tab.subn[0][i] : (r == 1 ? A.tab.subn[1][i] : A.tab.subn[2][i]);
                    double G[NV];
                    ec_flux<DIM, true>(r, si, qa, qb, s1, s2, P, quirk3d, G);
                    const double fa = alpha_e * A.tab.fvs[i];
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        acc[i][v] += fa * (gprev[v] - G[v]);
                        gprev[v] = G[v];
                    }
                }
                const double fl = alpha_e * A.tab.fvs[NP - 1];
#pragma unroll
                for (int v = 0; v < NV; v++) acc[NP - 1][v] += fl * gprev[v];
            }
        }
    }

    // Every start load has been used by now on the lanes that have a line;
    // saying so (vmcnt 0, which costs nothing here) keeps lanes without one
    // from leaving those loads pending in the compiler's count, which would
    // make it drain the assembly loads below one by one as they are issued.
    // (The immediate is gfx9's s_waitcnt encoding: vmcnt 0, expcnt 7,
    // lgkmcnt 15.)
    __builtin_amdgcn_s_waitcnt(0x0F70);

    // ---- the assembly's inputs (viscous volume term, u0), issued before
    // the direction sums so their latency overlaps them ----
    double vsp[NPASS][NV - 1], u0p[NPASS][NV];
#pragma unroll
    for (int p = 0; p < NPASS; p++) {
        const int i = t + p * BLOCK;
        const size_t gi = (size_t)ebase * K + i;
#pragma unroll
        for (int v = 0; v < NV - 1; v++)
            vsp[p][v] = (A.parabolic && i < nact) ? (double)static_cast<const VT *>(A.visc)[v * NK + gi] : 0.0;
#pragma unroll
        for (int v = 0; v < NV; v++) u0p[p][v] = (A.mode == 2 && i < nact) ? A.u0[v * NK + gi] : 0.0;
    }

    // ---- node sums over the directions, in direction order ----
    if (!BLEND) __syncthreads();   // sP overwrites sQ
#pragma unroll
    for (int rr = 0; rr < DIM; rr++) {
        if (lact && r == rr) {
#pragma unroll
            for (int a = 0; a < NP; a++) {
                const int nd = lnode(a);
#pragma unroll
                for (int v = 0; v < NV; v++) sP(el, v, nd) = (rr == 0) ? acc[a][v] : sP(el, v, nd) + acc[a][v];
            }
        }
        __syncthreads();
    }

    // ---- viscous volume term, stage update, store ----
#pragma unroll
    for (int p = 0; p < NPASS; p++) {
        const int i = t + p * BLOCK;
        if (i >= nact) continue;
        const int e2 = i / K, n = i - e2 * K;
        const size_t gi = (size_t)ebase * K + i;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            double tot = sP(e2, v, node_slot<NP>(n));
            if (A.parabolic && v > 0) tot += vsp[p][v - 1] * A.geo.invJ;
            const double yv = sU(e2, v, node_slot<NP>(n));
            double o;
            if (A.mode == 0) o = tot;
            else if (A.mode == 1) o = yv + dt * tot;
            else o = A.c0 * u0p[p][v] + A.c1 * (yv + dt * tot);
            A.out[v * NK + gi] = o;
        }
    }
    block_lam_max<BLOCK>(lam, lam_old, A.lam, sRed, blockIdx.x);
}

} // namespace prandtl
