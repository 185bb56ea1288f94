// Host-side launch templates of k_visc and k_rhs and the context they read.
// Each (dimension, order) pair is instantiated in its own translation unit
// (inst_*.hip, listed by PRANDTL_ORDERS) so the library builds in parallel.
#pragma once
#include <string>
#include <vector>
#include "dg_device.h"
#include "dg_line.h"
#include "dg_visc_mx.h"

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
    // time (TimeDependentOperator::t): host mirror, device copy, the device
    // time a stage reads (the loop's while prandtl_run launches) and the
    // stage's offset factor (0, 1, 1/2 of dt in an RK3 step)
    double t = 0.0;
    double *d_time = nullptr;
    uint64_t geom_fp = 0;   // geometry fingerprint kept in checkpoints
    const double *tp = nullptr;
    double stage_tc = 0.0;
    double *d_bfn = nullptr;       // split outer states per attribute (quadrilateral meshes)
    int nattr = 0;
    double graph_dt = 0.0;        // value last copied to d_dt
    bool graph_dt_set = false;
    int rhs_algo = 0;             // flux-differencing kernel: PRANDTL_ALGO_TILE (k_rhs) or _LINE (k_rhs_ln)
    int rhs_shape = 0;            // k_rhs launch shape (only shapes rhs_shape_ok accepts are set)
    int visc_shape = 0;           // k_visc launch shape (visc_shape_ok)
    // device time loop (prandtl_run)
    LoopState *d_loop = nullptr;
    const int *skip = nullptr;    // set while launching the loop's stages
};

namespace prandtl {

// message of the last failed create (no context to hold it)
inline std::string g_create_err;

inline int fail(prandtl_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

inline int hip_check(prandtl_ctx *c, hipError_t e, const char *what)
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

// The MFMA viscous kernel serves the mixed form of the line algorithm on
// periodic 3-D boxes at N = 3.
inline bool visc_mx_ok(const prandtl_ctx *c)
{
    return c->rhs_algo == PRANDTL_ALGO_LINE && c->prm.precision == PRANDTL_MIXED && c->dim == 3 && c->Np == 4 &&
           !c->gen && !c->axi && c->geo.bcmask == 0;
}

template <int DIM, int NP> int launch_visc_nd(prandtl_ctx *c, const double *y, double *grad, hipStream_t s)
{
    VArgs a;
    a.tab = c->tab; a.geo = c->geo; a.phys = c->phys;
    a.quirks = c->prm.quirks;
    a.y = y; a.visc = c->d_visc; a.fvf = c->d_fvf; a.grad = grad;
    a.skip = c->skip;
    ProfScope ps(c, s, 1);
    if constexpr (DIM == 3 && NP == 4) {
        if (visc_mx_ok(c)) {
            using Mc = MxCfg<NP>;
            const int nb = (c->NE + Mc::EPB - 1) / Mc::EPB;
            hipLaunchKernelGGL((k_visc_mx<NP>), dim3(nb), dim3(Mc::BLOCK), 0, s, a);
            return hip_check(c, hipGetLastError(), "k_visc_mx launch");
        }
    }
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

// The line kernel serves periodic Cartesian boxes of the orders line_built
// lists.
inline bool line_ok(const prandtl_ctx *c)
{
    return line_built(c->dim, c->Np) && !c->gen && !c->axi && c->geo.bcmask == 0;
}

// The default algorithm where the line kernel exists: it is the faster one
// at N = 2 and 3, and with blending at N = 4; unblended N = 4 runs the tile
// kernel faster (12^3 boxes on MI355X: Euler 35.6 against 42.0 us per RHS,
// Navier-Stokes 54.9 against 62.3).
inline bool line_default(const prandtl_ctx *c)
{
    return line_ok(c) && (c->Np <= 4 || c->prm.blending);
}

template <int DIM, int NP, bool BLEND>
int launch_rhs_t(prandtl_ctx *c, const KArgs &a, hipStream_t s)
{
    ProfScope ps(c, s, 0);
    const bool qns3 = DIM == 3 && c->prm.parabolic && c->prm.quirks;
    if constexpr (line_built(DIM, NP)) {
        if (c->rhs_algo == PRANDTL_ALGO_LINE && line_ok(c)) {
            using Lc = LCfg<DIM, NP>;
            const int nb = (c->NE + Lc::EPB - 1) / Lc::EPB;
            if (qns3)
                hipLaunchKernelGGL((k_rhs_ln<DIM, NP, BLEND, true, double>), dim3(nb), dim3(Lc::BLOCK), 0, s, a);
            else if (c->prm.parabolic && c->prm.precision == PRANDTL_MIXED)
                hipLaunchKernelGGL((k_rhs_ln<DIM, NP, BLEND, false, float>), dim3(nb), dim3(Lc::BLOCK), 0, s, a);
            else hipLaunchKernelGGL((k_rhs_ln<DIM, NP, BLEND, false, double>), dim3(nb), dim3(Lc::BLOCK), 0, s, a);
            return hip_check(c, hipGetLastError(), "k_rhs_ln launch");
        }
    }
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

inline KArgs make_kargs(const prandtl_ctx *c, const double *y, const double *u0, double *out, int mode,
                 double dt, double c0, double c1, const double *dtp)
{
    KArgs a;
    a.tab = c->tab; a.geo = c->geo; a.phys = c->phys;
    a.blend = c->prm.blending; a.parabolic = c->prm.parabolic;
    a.ec = c->prm.ec_surface; a.quirks = c->prm.quirks;
    a.mode = mode; a.dt = dt; a.c0 = c0; a.c1 = c1;
    a.dtp = dtp;
    a.tp = c->tp ? c->tp : c->d_time;
    a.tc = c->stage_tc;
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


// Every (dimension, order) the library is built for: X(DIM, NP).
#define PRANDTL_ORDERS(X) \
    X(1, 2) X(1, 3) X(1, 4) X(1, 5) X(1, 6) X(1, 7) X(1, 8) \
    X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(2, 6) \
    X(3, 3) X(3, 4) X(3, 5) X(3, 6) X(3, 7) X(3, 8)

// The per-order entry points the host unit calls (explicitly instantiated
// in inst_*.hip; "extern" there is empty).
#define PRANDTL_INSTANTIATE(EXT, D, N) \
    EXT template int stage_t<D, N>(prandtl_ctx *, const double *, const double *, double *, int, double, double, \
                                   double, hipStream_t); \
    EXT template int launch_visc<D, N>(prandtl_ctx *, const double *, double *, hipStream_t); \
    EXT template int launch_rhs<D, N>(prandtl_ctx *, const double *, const double *, double *, int, double, double, \
                                      double, hipStream_t);

} // namespace prandtl
