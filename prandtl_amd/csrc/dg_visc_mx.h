// k_visc_mx: k_visc of the mixed-precision form on periodic 3-D boxes at
// N = 3, with its two families of (N+1)x(N+1) derivative contractions on the
// FP32 matrix cores (gfx950 v_mfma_f32_4x4x1f32).
//
// Same operator as k_visc<3,4,MIXED=true> (LiftingBR1.cpp:10-66,156-219;
// DGSEMOperator.cpp:82-159; NavierStokesFlux.cpp:12-58; DGSEMIntegrator.cpp:
// 617-641,705-738): entropy variables in FP64, shifted by their value at the
// element's node 0 and rounded to FP32; everything after in FP32.  Phases:
//
//  A (one node per thread)  w(u) of the element's nodes and of the
//     neighbours' face traces into LDS;
//  B (one line task per thread: direction r, line l, variable v)  the BR1
//     volume derivative of w along the line, sum_m D_am w_m for its NP = 4
//     nodes, as four MFMA steps: lane 4b+i supplies D[i][m] as A and every lane
//     its own line's w_m as B, so each lane gets back its own line's four
//     derivatives (the instruction's 16 blocks of 4x4x1 products); then the
//     scaling by the metric and the face jumps at the line's two ends;
//  C (one node per thread)  gradient conversion, viscous flux, contravariant
//     fluxes into LDS, the face traces Fv.n to HBM;
//  D (line tasks: direction r, line l, flux variable v)  the weak-form
//     divergence sum_m D^_am (Ja Fv)_m by the same four MFMA steps;
//  E (one node per thread)  the sum over directions, written as FP32.
//
// Line lanes hold one line each (192 of the 256 threads at N = 3); the MFMA
// runs with the whole wave, and lanes past the last element supply zeros
// whose outputs are dropped.
#pragma once
#include "dg_device.h"

namespace prandtl {

typedef float mx_v4f __attribute__((ext_vector_type(4)));

template <int NP> struct MxCfg {
    static constexpr int DIM = 3, NV = 5;
    static constexpr int K = NP * NP * NP, NFP = NP * NP;
    static constexpr int EPB = 256 / K;
    static constexpr int BLOCK = EPB * K;
    static constexpr int KP = K + K / NP;   // padded node rows, as in k_rhs_ln
};

// sum_m M[i][m] x_m for the lane's own line x (NP = 4), M given per lane as
// its row i = lane % 4 (mrow[m] = M[lane % 4][m]): out[a] = sum_m M[a][m] x_m
__device__ __forceinline__ mx_v4f mx_line4(const float *mrow, const float *x)
{
    mx_v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 4; m++) c = __builtin_amdgcn_mfma_f32_4x4x1f32(mrow[m], x[m], c, 0, 0, 0);
    return c;
}

template <int NP>
__global__ void __launch_bounds__((MxCfg<NP>::BLOCK)) k_visc_mx(const VArgs A)
{
    using C = MxCfg<NP>;
    constexpr int DIM = C::DIM, NV = C::NV, K = C::K, NFP = C::NFP, EPB = C::EPB, BLOCK = C::BLOCK, KP = C::KP;
    static_assert(NP == 4, "4x4x1 MFMA blocks: N = 3");
    using R = float;
    if (A.skip && *A.skip) return;   // uniform: before any barrier
    // region X: entropy offsets W and neighbour traces WN, later the
    // contravariant fluxes FV; region Y: directional derivatives G, later the
    // divergence partials VP
    constexpr int XW = EPB * NV * KP, XWN = EPB * 2 * DIM * NV * NFP, XFV = EPB * DIM * (NV - 1) * KP;
    constexpr int XS = (XW + XWN) > XFV ? XW + XWN : XFV;
    constexpr int YG = EPB * DIM * NV * KP, YVP = EPB * DIM * (NV - 1) * KP;
    constexpr int YS = YG > YVP ? YG : YVP;
    __shared__ R sX[XS];
    __shared__ R sY[YS];
    __shared__ double sUl[EPB][NV];   // state of node K-1 (the faithful stale state)
    __shared__ double sWref[EPB][NV];
    auto sW = reinterpret_cast<R (*)[NV][KP]>(sX);
    auto sWN = reinterpret_cast<R (*)[2 * DIM][NV][NFP]>(sX + XW);
    auto sFV = reinterpret_cast<R (*)[DIM][NV - 1][KP]>(sX);
    auto sG = reinterpret_cast<R (*)[DIM][NV][KP]>(sY);
    auto sVP = reinterpret_cast<R (*)[DIM][NV - 1][KP]>(sY);
    auto pad = [](int n) { return n + n / NP; };

    const int t = threadIdx.x;
    const Phys &P = A.phys;
    const int NE = A.geo.NE;
    const size_t NK = (size_t)NE * K;
    const int ebase = xcd_remap(blockIdx.x, gridDim.x) * EPB;
    const int el = t / K;
    const int n = t - el * K;
    const int e = ebase + el;
    const bool act = e < NE;
    const int ii[3] = {n % NP, (n / NP) % NP, n / (NP * NP)};
    R *visc = reinterpret_cast<R *>(A.visc);
    R *fvf = reinterpret_cast<R *>(A.fvf);
    const R gm1 = (R)P.gm1, gm1i = (R)P.gm1i;
    const ViscCoef<R> vcoef{(R)P.mu, (R)P.lam_h, P.sutherland != 0, (R)P.R_gas, (R)P.gamma, (R)P.PrInv};
    const R nrm[3] = {(R)A.geo.nrm[0], (R)A.geo.nrm[1], (R)A.geo.nrm[2]};
    const R invJ = (R)A.geo.invJ, aface = (R)A.geo.aface, half = 0.5f;
    // this lane's rows of D and D^ for the MFMA A operand
    float drow[NP], dhrow[NP];
#pragma unroll
    for (int m = 0; m < NP; m++) {
        drow[m] = (float)A.tab.D[(t & 3) * MAXNP + m];
        dhrow[m] = (float)A.tab.Dh[(t & 3) * MAXNP + m];
    }

    // ---- phase A: states, entropy variables, neighbour traces ----
    constexpr int NJ = EPB * 2 * DIM * NFP;
    constexpr int JT = (NJ + BLOCK - 1) / BLOCK;
    double ys[NV], wv[NV];
    double tv[JT][NV];
    if (act)
#pragma unroll
        for (int v = 0; v < NV; v++) ys[v] = A.y[v * NK + (size_t)e * K + n];
#pragma unroll
    for (int k = 0; k < JT; k++) {
        const int j = t + k * BLOCK;
        const int ej = j / (2 * DIM * NFP);
        const int eo = ebase + ej;
        if (j < NJ && eo < NE) {
            const int rem = j - ej * (2 * DIM * NFP);
            const int f = rem / NFP, q = rem - f * NFP;
            const int r = f >> 1, side = f & 1;
            const int en = neighbor_lane(A.geo, eo, r, side);
            const int nn = line_node<DIM, NP>(r, q, side ? 0 : NP - 1);
#pragma unroll
            for (int v = 0; v < NV; v++) tv[k][v] = A.y[v * NK + (size_t)en * K + nn];
        }
    }
    if (act) {
        if (n == K - 1)
#pragma unroll
            for (int v = 0; v < NV; v++) sUl[el][v] = ys[v];
        cons2entropy<DIM>(ys, P, wv);
        if (n == 0)
#pragma unroll
            for (int v = 0; v < NV; v++) sWref[el][v] = wv[v];
    }
    __syncthreads();
    if (act)
#pragma unroll
        for (int v = 0; v < NV; v++) sW[el][v][pad(n)] = (R)(wv[v] - sWref[el][v]);
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
#pragma unroll
            for (int v = 0; v < NV; v++) sWN[ej][f][v][q] = (R)(w[v] - sWref[ej][v]);
        }
    }
    __syncthreads();

    // ---- phase B: BR1 derivatives along every line (MFMA), face jumps ----
    // lane t < EPB DIM NFP holds line (element lel, direction lr, line ll)
    // for every variable; the waves beyond them skip the line phases (whole
    // waves, so each MFMA runs with a full wave)
    constexpr int NL = EPB * DIM * NFP;
    const int lel = t / (DIM * NFP);
    const int lrl = t - lel * (DIM * NFP);
    const int lr = lrl / NFP, ll = lrl - lr * NFP;
    const bool lwave = (t & ~63) < NL;          // this wave has line lanes
    const bool live = t < NL && ebase + lel < NE;
    const int pb = lr == 0 ? ll * (NP + 1) : (lr == 1 ? (ll / NP) * NP * (NP + 1) + ll % NP : ll + ll / NP);
    const int ps = lr == 0 ? 1 : (lr == 1 ? NP + 1 : NP * (NP + 1));
    const R nr = lr == 0 ? nrm[0] : (lr == 1 ? nrm[1] : nrm[2]);
    if (lwave) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            float x[NP];
#pragma unroll
            for (int m = 0; m < NP; m++) x[m] = live ? sW[lel][v][pb + m * ps] : 0.0f;
            const mx_v4f c = mx_line4(drow, x);
            if (live) {
                // +1/2 (w2 - w1) n / (w0 J) on both sides (LiftingBR1.cpp:37-65)
                const R jl = half * (x[0] - sWN[lel][2 * lr][v][ll]);
                const R jh = half * (sWN[lel][2 * lr + 1][v][ll] - x[NP - 1]);
#pragma unroll
                for (int a = 0; a < NP; a++) {
                    R g = (c[a] * nr) * invJ;
                    if (a == 0) g += aface * (jl * nr);
                    if (a == NP - 1) g += aface * (jh * nr);
                    sG[lel][lr][v][pb + a * ps] = g;
                }
            }
        }
    }
    __syncthreads();

    // ---- phase C: gradient conversion, viscous flux, contravariant fluxes ----
    R fl[NV - 1][DIM];
    if (act) {
        R gp[DIM][NV];
        R yr[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) yr[v] = (R)ys[v];
#pragma unroll
        for (int c = 0; c < DIM; c++)
#pragma unroll
            for (int v = 0; v < NV; v++) gp[c][v] = sG[el][c][v][pad(n)];
#pragma unroll
        for (int c = 0; c < DIM; c++) entgrad2primgrad<DIM, R>(yr, gm1, gm1i, A.quirks != 0, gp[c]);
        if (A.grad)
#pragma unroll
            for (int c = 0; c < DIM; c++)
#pragma unroll
                for (int v = 0; v < NV; v++) A.grad[(c * NV + v) * NK + (size_t)e * K + n] = (double)gp[c][v];
        // volume viscous flux with the stale node K-1 state (DGSEMIntegrator.cpp:620)
        R sv[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) sv[v] = A.quirks ? (R)sUl[el][v] : yr[v];
        viscous_flux<DIM, R>(sv, gp, gm1, gm1i, vcoef, fl);
        bool onface = false;
#pragma unroll
        for (int c = 0; c < DIM; c++) onface |= (ii[c] == 0 || ii[c] == NP - 1);
        // the contravariant volume fluxes go to LDS after the barrier (they
        // overwrite W); face traces Fv.n with the node's own state now
        if (onface) {
            R ff[NV - 1][DIM];
            if (A.quirks) viscous_flux<DIM, R>(yr, gp, gm1, gm1i, vcoef, ff);
            else
#pragma unroll
                for (int v = 0; v < NV - 1; v++)
#pragma unroll
                    for (int c = 0; c < DIM; c++) ff[v][c] = fl[v][c];
#pragma unroll
            for (int c = 0; c < DIM; c++) {
                if (ii[c] == 0 || ii[c] == NP - 1) {
                    const int side = (ii[c] == NP - 1);
                    const int l = line_of<DIM, NP>(c, ii[0], ii[1], ii[2]);
                    R *dst = fvf + (((size_t)e * 2 * DIM + 2 * c + side) * (NV - 1)) * NFP;
#pragma unroll
                    for (int v = 1; v < NV; v++) dst[(v - 1) * NFP + l] = ff[v - 1][c];
                }
            }
        }
    }
    __syncthreads();   // W and G are dead: FV overwrites W, VP overwrites G
    if (act)
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int v = 1; v < NV; v++) sFV[el][r][v - 1][pad(n)] = nrm[r] * fl[v - 1][r];
    __syncthreads();

    // ---- phase D: weak-form viscous divergence along every line (MFMA) ----
    if (lwave) {
#pragma unroll
        for (int v = 0; v < NV - 1; v++) {
            float x[NP];
#pragma unroll
            for (int m = 0; m < NP; m++) x[m] = live ? sFV[lel][lr][v][pb + m * ps] : 0.0f;
            const mx_v4f c = mx_line4(dhrow, x);
            if (live)
#pragma unroll
                for (int a = 0; a < NP; a++) sVP[lel][lr][v][pb + a * ps] = c[a];
        }
    }
    __syncthreads();

    // ---- phase E: sum over directions ----
    if (act) {
#pragma unroll
        for (int v = 0; v < NV - 1; v++) {
            R s = sVP[el][0][v][pad(n)];
#pragma unroll
            for (int r = 1; r < DIM; r++) s += sVP[el][r][v][pad(n)];
            visc[v * NK + (size_t)e * K + n] = s;
        }
    }
}

} // namespace prandtl
