// Host unit of libprandtl_hip.so: the C ABI of include/prandtl_hip.h
// (contexts, RHS and RK3 entry points, hipGraph capture, autotuner, device
// time loop, checkpoints, primitives) plus the small utility kernels and
// k_ens, the batched-ensemble kernel.
//
// The DGSEM kernels themselves, k_visc and k_rhs, are templates in
// dg_device.h; their launch code (dg_launch.h) is instantiated per
// (dimension, order) in inst.hip, and this unit calls those instances
// through the PRANDTL_DISPATCH switch below.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/prandtl_hip.h"
#include "dg_launch.h"
#include "mesh.h"

namespace prandtl {

// ---------------------------------------------------------------------------
// k_ens: one launch per RK stage over the element tiles of an ensemble of
// independent 2-D Euler cases with mixed polynomial degree and blending.
//
// Workgroup b takes tile b of a host-built list sorted by decreasing cost; the
// workgroup dispatcher hands workgroups to CUs in index order as slots free,
// so the list is the work queue and needs no device counter.  k_ens_q below
// is the persistent form over the same list; prandtl_ensemble_autotune
// times both.
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
__device__ __forceinline__ void ens_tile(const KArgs &A, int g, double *smem, int tid = -1)
{
    rhs_body<2, NP, BLEND, false, double, ENS_BLOCK>(A, g, g, *reinterpret_cast<RhsSmem<2, NP, BLEND, false, ENS_BLOCK> *>(smem),
                                                     tid);
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

// Persistent form: a fixed grid (a chosen number of workgroups per CU) loops
// over the same cost-sorted tile list, each workgroup taking the next tile
// from a device counter until the list is exhausted (every workgroup leaves
// the loop when the counter passes the list's end; no workgroup waits on
// another).  Arithmetic is that of k_ens, so results are bitwise the same.
//
// Two details keep the loop a plain uniform loop in the generated code:
//  - the thread index is passed through an empty asm once per iteration, and
//    both the fetch (tid == 0) and the tile body use that value.  With
//    threadIdx.x itself the compiler threads the body's final thread-0 branch
//    (the wave-speed slot update) straight into the next iteration's fetch;
//    the back edge then splits wave 0 (lane 0 to the atomic, lanes 1..63 to
//    the barrier), and the lanes that skip the fetch re-read a stale tile
//    index forever.  The opaque value also keeps each variant's index
//    arithmetic inside the loop instead of hoisted and live across it;
//  - the tile index read back from LDS goes through readfirstlane, so the
//    tile's fields, the variant branch and the case's arguments are scalar
//    as in k_ens.
__global__ void __launch_bounds__(ENS_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_ens_q(const KArgs *__restrict__ cases, const EnsTile *__restrict__ tiles, int ntiles, int *counter)
{
    __shared__ __align__(16) double smem[(ENS_SMEM + 7) / 8];
    __shared__ int s_tile;
    for (;;) {
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (tid == 0) s_tile = atomicAdd(counter, 1);
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();   // every thread has read s_tile before the next fetch
        if (i >= ntiles) break;
        const EnsTile T = tiles[i];
        const KArgs &A = cases[T.c];
        switch (T.v) {
        case 0: ens_tile<3, false>(A, T.g, smem, tid); break;
        case 1: ens_tile<3, true>(A, T.g, smem, tid); break;
        case 2: ens_tile<4, false>(A, T.g, smem, tid); break;
        case 3: ens_tile<4, true>(A, T.g, smem, tid); break;
        case 4: ens_tile<5, false>(A, T.g, smem, tid); break;
        case 5: ens_tile<5, true>(A, T.g, smem, tid); break;
        case 6: ens_tile<6, false>(A, T.g, smem, tid); break;
        case 7: ens_tile<6, true>(A, T.g, smem, tid); break;
        default: break;
        }
    }
}

// the three stage counters to zero (a kernel, not a memset: it sits in the
// step's graph)
__global__ void k_zero_counters(int *c)
{
    if (threadIdx.x < 3) c[threadIdx.x] = 0;
}

constexpr int ENS_MAX_CASES = 256;
struct DtArgs { double dt[ENS_MAX_CASES]; };
__global__ void k_set_dts(double *dst, int n, const DtArgs a)
{
    const int i = threadIdx.x;
    if (i < n) dst[i] = a.dt[i];
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
// t += dt after an RK3 step (dt from device memory in graph replays)
__global__ void k_time_advance(double *t, const double *dtp, double dt, const int *skip)
{
    if (skip && *skip) return;
    *t += dtp ? *dtp : dt;
}
__global__ void k_reset_int(int *p) { *p = 0; }



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

// the per-order launch templates live in their own units (inst.hip)
namespace prandtl {
#define PRANDTL_EXTERN(D, N) PRANDTL_INSTANTIATE(extern, D, N)
PRANDTL_ORDERS(PRANDTL_EXTERN)
#undef PRANDTL_EXTERN
} // namespace prandtl

namespace {

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
        return c->prm.parabolic &&
               (shape == 0 || (!c->geo.bcmask && !c->gen && !visc_mx_ok(c) && visc_shape_ok(c->dim, c->Np, shape)));
    if (kernel != PRANDTL_KERNEL_RHS) return false;
    if (shape == 0) return true;
    if (c->rhs_algo == PRANDTL_ALGO_LINE) return false;
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
//
// host_ctx (boxes only): fill this caller-owned context's tables and stop
// before any device memory is allocated (prandtl_rhs_noalloc).
// FNV-1a, 64 bit, continuing from h (checkpoint checksums and mesh fingerprints)
uint64_t fnv1a64(const void *data, size_t n, uint64_t h = 1469598103934665603ull)
{
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Fingerprint of a case's geometry, kept in checkpoints so that a state is
// loaded only into the mesh it was written for: a box's element counts and
// extents, or an unstructured mesh's vertex coordinates, connectivity and
// per-element corners (order and variable count are in the header itself).
uint64_t geom_fingerprint(const prandtl_mesh_desc *m, const prandtl_quad_mesh *qm, const prandtl_hex_mesh *hm)
{
    if (qm || hm) {
        const int nvert = qm ? qm->nvert : hm->nvert, ne = qm ? qm->nelem : hm->nelem;
        const int sd = qm ? 2 : 3, nc = qm ? 4 : 8;
        const double *vert = qm ? qm->vert : hm->vert, *ec = qm ? qm->elem_coords : hm->elem_coords;
        const int *elem = qm ? qm->elem : hm->elem;
        uint64_t h = fnv1a64(&sd, sizeof(sd));
        h = fnv1a64(&nvert, sizeof(nvert), h);
        h = fnv1a64(vert, sizeof(double) * sd * (size_t)nvert, h);
        h = fnv1a64(&ne, sizeof(ne), h);
        h = fnv1a64(elem, sizeof(int) * nc * (size_t)ne, h);
        if (ec) h = fnv1a64(ec, sizeof(double) * nc * sd * (size_t)ne, h);
        return h;
    }
    uint64_t h = fnv1a64(&m->dim, sizeof(m->dim));
    for (int r = 0; r < m->dim && r < 3; r++) {
        h = fnv1a64(&m->nelem[r], sizeof(int), h);
        h = fnv1a64(&m->lo[r], sizeof(double), h);
        h = fnv1a64(&m->hi[r], sizeof(double), h);
    }
    return h;
}

int create_impl(const prandtl_mesh_desc *mesh, const prandtl_quad_mesh *qm, const prandtl_params *prm,
                prandtl_ctx **out, const prandtl_hex_mesh *hm = nullptr, prandtl_ctx *host_ctx = nullptr)
{
    if (!mesh || !prm || !out) return fail(nullptr, PRANDTL_ERR_ARG, "null argument");
    if (host_ctx && (qm || hm)) return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "host-only set-up is for boxes");
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
    prandtl_ctx *c = host_ctx ? host_ctx : new (std::nothrow) prandtl_ctx();
    if (!c) return fail(nullptr, PRANDTL_ERR_NOMEM, "host allocation");
    c->dim = dim; c->order = N; c->Np = N + 1; c->nv = dim + 2;
    c->mesh = *mesh;
    c->geom_fp = geom_fingerprint(mesh, qm, hm);
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
        if (!host_ctx) delete c;
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
            if (!host_ctx) delete c;
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
            if (!host_ctx) delete c;
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
    c->geo.inv_n0 = 1.0 / c->geo.n[0];
    c->geo.inv_n01 = 1.0 / ((double)c->geo.n[0] * c->geo.n[1]);
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

    if (host_ctx) return PRANDTL_OK;   // tables only: the caller brings device memory
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
    if (e == hipSuccess) e = hipMalloc(&c->d_time, sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->d_time, 0, sizeof(double));
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
        constexpr int NPART = 14;
        const size_t parts[NPART] = {bytes(qg.met), bytes(qg.jac), bytes(qg.fnor), bytes(qg.subn), bytes(qstates),
                                     bytes(qg.fface), bytes(qg.fnb), bytes(qg.fnf), bytes(qg.rnode),
                                     bytes(qg.axm), bytes(qg.axj), bytes(qg.axr), bytes(qg.corner), bytes(qg.fxy)};
        size_t off[NPART + 1] = {0};
        for (int i = 0; i < NPART; i++) off[i + 1] = off[i] + ((parts[i] + 255) / 256) * 256;
        e = hipMalloc(&c->d_geo, off[NPART]);
        const void *src[NPART] = {qg.met.data(), qg.jac.data(), qg.fnor.data(), qg.subn.data(), qstates.data(),
                                  qg.fface.data(), qg.fnb.data(), qg.fnf.data(), qg.rnode.data(),
                                  qg.axm.data(), qg.axj.data(), qg.axr.data(), qg.corner.data(), qg.fxy.data()};
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
        if (qm) c->geo.fxy = reinterpret_cast<const double *>(b + off[13]);
        c->nattr = qm ? qm->nattr : hm->nattr;
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
    // the line kernel where it serves this case
    c->rhs_algo = line_default(c) ? PRANDTL_ALGO_LINE : PRANDTL_ALGO_TILE;
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
    if (c->d_time) (void)hipFree(c->d_time);
    if (c->d_bfn) (void)hipFree(c->d_bfn);
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

} // extern "C"

namespace {

// Device scratch of prandtl_rhs_noalloc: one copy of this array per GPU the
// library runs on, held until the process ends.  Host code never uses the
// name itself as an address (it is a host-side stand-in there): each
// device's address comes from hipGetSymbolAddress, looked up once per device
// with that device current, and only that address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, PRANDTL_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(nullptr, PRANDTL_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, PRANDTL_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PRANDTL_OK;
}

struct BoxStates {
    double v[30];
};

// Clears the running wave-speed slots and writes the boundary states (a
// kernel, not a memset, so the call may sit in a captured graph).
__global__ void k_noalloc_init(double *lam, int nlam, double *bcs, BoxStates b, int nbcs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nlam) lam[i] = 0.0;
    if (i < nbcs) bcs[i] = b.v[i];
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

} // namespace

extern "C" {

int prandtl_rhs_noalloc(const prandtl_mesh_desc *mesh, const prandtl_params *params, const double *u,
                        double *dudt, void *stream)
{
    if (!mesh || !params || !u || !dudt) return fail(nullptr, PRANDTL_ERR_ARG, "null argument");
    if (u == dudt) return fail(nullptr, PRANDTL_ERR_ARG, "u and dudt alias");
    prandtl_ctx c;
    prandtl_ctx *unused = nullptr;
    int rc = create_impl(mesh, nullptr, params, &unused, nullptr, &c);
    if (rc) return rc;
    // scratch layout: wave-speed slots, alpha, eta, box states, viscous volume
    // sums and face traces, each 256-byte aligned
    const size_t NK = (size_t)c.NE * c.K;
    const size_t sizes[6] = {sizeof(double) * c.NE, sizeof(double) * c.NE, sizeof(double) * c.NE,
                             sizeof(double) * 30,
                             c.prm.parabolic ? sizeof(double) * (c.nv - 1) * NK : 0,
                             c.prm.parabolic ? sizeof(double) * (size_t)c.NE * 2 * c.dim * (c.nv - 1) * c.Nfp : 0};
    size_t off[7] = {0};
    for (int i = 0; i < 6; i++) off[i + 1] = off[i] + align256(sizes[i]);
    if (off[6] > NOALLOC_SCRATCH_BYTES) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "the case needs %zu bytes of scratch; prandtl_rhs_noalloc holds %zu",
                      off[6], NOALLOC_SCRATCH_BYTES);
        return fail(nullptr, PRANDTL_ERR_ARG, msg);
    }
    void *base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;
    char *b = static_cast<char *>(base);
    c.d_lam = reinterpret_cast<double *>(b + off[0]);
    c.d_alpha = reinterpret_cast<double *>(b + off[1]);
    c.d_eta = reinterpret_cast<double *>(b + off[2]);
    c.d_bcs = reinterpret_cast<double *>(b + off[3]);
    c.d_visc = c.prm.parabolic ? reinterpret_cast<double *>(b + off[4]) : nullptr;
    c.d_fvf = c.prm.parabolic ? reinterpret_cast<double *>(b + off[5]) : nullptr;
    c.geo.bcs = c.geo.bcmask ? c.d_bcs : nullptr;
    c.rhs_algo = line_default(&c) ? PRANDTL_ALGO_LINE : PRANDTL_ALGO_TILE;
    BoxStates bs;
    std::memcpy(bs.v, mesh->bc_state, sizeof(bs.v));
    hipStream_t s = (hipStream_t)stream;
    const int n = c.NE > 30 ? c.NE : 30;
    hipLaunchKernelGGL(k_noalloc_init, dim3((n + 255) / 256), dim3(256), 0, s, c.d_lam, c.NE, c.d_bcs, bs,
                       c.geo.bcmask ? 30 : 0);
    rc = hip_check(&c, hipGetLastError(), "k_noalloc_init launch");
    if (!rc) rc = stage(&c, u, nullptr, dudt, 0, 0.0, 0.0, 0.0, s);
    if (rc) g_create_err = c.err;
    return rc;
}

} // extern "C"

namespace {

static int rk3_stages(prandtl_ctx *c, double *u, double dt, hipStream_t s)
{
    // y1 = u + dt L(u); y2 = 3/4 u + 1/4 (y1 + dt L(y1)); u = 1/3 u + 2/3 (y2 + dt L(y2)),
    // evaluated at t, t + dt and t + dt/2 (RK3SSPSolver's SetTime calls)
    c->stage_tc = 0.0;
    int rc = stage(c, u, nullptr, c->d_ya, 1, dt, 0.0, 0.0, s);
    c->stage_tc = 1.0;
    if (!rc) rc = stage(c, c->d_ya, u, c->d_yb, 2, dt, 3.0 / 4, 1.0 / 4, s);
    c->stage_tc = 0.5;
    if (!rc) rc = stage(c, c->d_yb, u, u, 2, dt, 1.0 / 3, 2.0 / 3, s);
    c->stage_tc = 0.0;
    if (!rc && !c->tp && c->d_bfn) {
        // t += dt on the device, for the time-dependent boundary states that
        // read it (the device loop advances its own time; without such
        // states only the host copy is kept, which saves a launch per step)
        hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(1), 0, s, c->d_time, c->dtp, dt, c->skip);
        rc = hip_check(c, hipGetLastError(), "time advance");
    }
    return rc;
}

} // namespace

extern "C" {

int prandtl_rk3_step(prandtl_ctx *c, double *u, double dt, void *stream)
{
    if (!c || !u) return fail(c, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    c->t += dt;   // the host mirror of the device time the step advances
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

int prandtl_set_time(prandtl_ctx *c, double t, void *stream)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (!std::isfinite(t)) return fail(c, PRANDTL_ERR_ARG, "time must be finite");
    c->t = t;
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, (hipStream_t)stream, c->d_time, t);
    return hip_check(c, hipGetLastError(), "set time");
}

int prandtl_get_time(const prandtl_ctx *c, double *t)
{
    if (!c || !t) return PRANDTL_ERR_ARG;
    *t = c->t;
    return PRANDTL_OK;
}

int prandtl_set_boundary_split(prandtl_ctx *c, int attr, const double *normal, double offset, double speed,
                               const double *below, const double *above)
{
    if (!c || !normal || !below || !above) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (!c->gen || c->dim != 2 || c->axi || c->prm.parabolic)
        return fail(c, PRANDTL_ERR_UNSUPPORTED,
                    "time-dependent outer states are built for 2-D quadrilateral meshes, Euler, non-axisymmetric");
    if (attr < 1 || attr > c->nattr) return fail(c, PRANDTL_ERR_ARG, "boundary attribute out of range");
    for (const double *g : {below, above}) {
        const double ke = 0.5 * (g[1] * g[1] + g[2] * g[2]) / g[0];
        const double pg = (c->prm.gamma - 1.0) * (g[3] - ke);
        if (!(g[0] > 0.0) || !(pg > 0.0) || !std::isfinite(pg))
            return fail(c, PRANDTL_ERR_ARG, "boundary state needs positive density and pressure");
    }
    hipError_t e = hipSuccess;
    if (!c->d_bfn) {
        // the device time has not been advanced by steps until now: bring it
        // to the host copy once every stream's work on it is done
        e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(c->d_time, &c->t, sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&c->d_bfn, sizeof(double) * 16 * c->nattr);
        if (e == hipSuccess) e = hipMemset(c->d_bfn, 0, sizeof(double) * 16 * c->nattr);
        if (e != hipSuccess) return hip_check(c, e, "boundary state table");
    }
    double f[16] = {1.0, normal[0], normal[1], 0.0, offset, speed};
    for (int v = 0; v < 4; v++) { f[6 + v] = below[v]; f[11 + v] = above[v]; }
    e = hipMemcpy(c->d_bfn + 16 * (size_t)(attr - 1), f, sizeof(f), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_check(c, e, "boundary state table");
    c->geo.bfn = c->d_bfn;
    drop_graph(c);
    return PRANDTL_OK;
}

int prandtl_set_algorithm(prandtl_ctx *c, int kernel, int algo)
{
    if (!c) return fail(c, PRANDTL_ERR_ARG, "null argument");
    if (kernel != PRANDTL_KERNEL_RHS || (algo != PRANDTL_ALGO_TILE && algo != PRANDTL_ALGO_LINE))
        return fail(c, PRANDTL_ERR_ARG, "unknown kernel or algorithm");
    if (algo == PRANDTL_ALGO_LINE && !line_ok(c))
        return fail(c, PRANDTL_ERR_UNSUPPORTED, "the line kernel serves periodic 3-D boxes of order 2..4");
    if (algo != c->rhs_algo) {
        drop_graph(c);
        c->rhs_algo = algo;
        c->rhs_shape = 0;
        c->visc_shape = 0;
    }
    return PRANDTL_OK;
}

int prandtl_get_algorithm(const prandtl_ctx *c, int kernel, int *algo)
{
    if (!c || !algo || kernel != PRANDTL_KERNEL_RHS) return PRANDTL_ERR_ARG;
    *algo = c->rhs_algo;
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
        c->tp = &c->d_loop->t;
        rc = stage(c, u, nullptr, c->d_yb, 0, 0.0, 0.0, 0.0, s);
        c->tp = nullptr;
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
        c->tp = &c->d_loop->t;
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
        c->tp = nullptr;
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
    // the operator's time follows the loop's
    c->t = L.t;
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, s, c->d_time, L.t);
    rc = hip_check(c, hipGetLastError(), "set time");
    if (rc) return rc;
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
    uint64_t geom;   // geom_fingerprint of the mesh the state was written on
};
constexpr char CHK_MAGIC[8] = {'P', 'R', 'D', 'T', 'L', 'C', 'H', 'K'};
constexpr int32_t CHK_VERSION = 2;

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

int chk_write(const char *dir, int cycle, double t, const prandtl_mesh_desc *mesh, int nv, const double *u,
              uint64_t geom)
{
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
    h.geom = geom;
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

int chk_read(const char *dir, int cycle, const prandtl_mesh_desc *mesh, int nv, double *u, double *t_out,
             uint64_t geom)
{
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
    if (h.geom != geom) return io_fail("checkpoint was written on another mesh (geometry fingerprint differs): " + path);
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

} // namespace

extern "C" {

int prandtl_checkpoint_write(const char *dir, int cycle, double t, const prandtl_mesh_desc *mesh, int nv,
                             const double *u)
{
    if (!dir || !mesh || !u || cycle < 0 || nv < 1 || mesh->dim < 1 || mesh->dim > 3)
        return fail(nullptr, PRANDTL_ERR_ARG, "bad arguments");
    return chk_write(dir, cycle, t, mesh, nv, u, geom_fingerprint(mesh, nullptr, nullptr));
}

int prandtl_checkpoint_read(const char *dir, int cycle, const prandtl_mesh_desc *mesh, int nv, double *u,
                            double *t_out)
{
    if (!dir || !mesh || !u || !t_out || cycle < 0 || nv < 1 || mesh->dim < 1 || mesh->dim > 3)
        return fail(nullptr, PRANDTL_ERR_ARG, "bad arguments");
    return chk_read(dir, cycle, mesh, nv, u, t_out, geom_fingerprint(mesh, nullptr, nullptr));
}

int prandtl_checkpoint_save(prandtl_ctx *c, const double *u, const char *dir, int cycle, double t, void *stream)
{
    if (!c || !u || !dir) return fail(c, PRANDTL_ERR_ARG, "null argument");
    std::vector<double> h((size_t)c->nv * c->NE * c->K);
    hipStream_t s = (hipStream_t)stream;
    int rc = hip_check(c, hipMemcpyAsync(h.data(), u, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s), "copy state");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(s), "sync");
    if (rc) return rc;
    rc = chk_write(dir, cycle, t, &c->mesh, c->nv, h.data(), c->geom_fp);
    if (rc) c->err = g_create_err;
    return rc;
}

int prandtl_checkpoint_load(prandtl_ctx *c, double *u, const char *dir, int cycle, double *t_out, void *stream)
{
    if (!c || !u || !dir || !t_out) return fail(c, PRANDTL_ERR_ARG, "null argument");
    std::vector<double> h((size_t)c->nv * c->NE * c->K);
    double t = 0.0;
    int rc = chk_read(dir, cycle, &c->mesh, c->nv, h.data(), &t, c->geom_fp);
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
    // launch form: 0 one workgroup per tile; g > 0 persistent, g workgroups per CU
    int mode = 0;
    int ncu = 0;
    int *d_counters = nullptr;   // [3] tile counters of the persistent form, one per stage
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

// one stage k of every case in the current launch form
void ens_stage(prandtl_ensemble *E, int k, int mode, hipStream_t s)
{
    if (mode == 0)
        hipLaunchKernelGGL(k_ens, dim3(E->ntiles), dim3(ENS_BLOCK), 0, s, E->d_args[k], E->d_tiles);
    else
        hipLaunchKernelGGL(k_ens_q, dim3(mode * E->ncu), dim3(ENS_BLOCK), 0, s, E->d_args[k], E->d_tiles, E->ntiles,
                           E->d_counters + k);
}

int ens_stages(prandtl_ensemble *E, hipStream_t s)
{
    if (E->mode > 0) hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, E->d_counters);
    for (int k = 0; k < 3; k++) ens_stage(E, k, E->mode, s);
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
    if (e == hipSuccess) e = hipMalloc(&E->d_counters, sizeof(int) * 3);
    if (e == hipSuccess) {
        int dev = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&E->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    }
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
    if (E->d_counters) (void)hipFree(E->d_counters);
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

int prandtl_ensemble_set_mode(prandtl_ensemble *E, int mode)
{
    if (!E || mode < 0 || mode > PRANDTL_ENS_MAX_PER_CU) return ens_fail(E, PRANDTL_ERR_ARG, "bad launch form");
    if (mode != E->mode) {
        if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
        if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
        E->mode = mode;
    }
    return PRANDTL_OK;
}

int prandtl_ensemble_autotune(prandtl_ensemble *E, double *const *u, const double *dt, int *mode_out, double *us_out,
                              void *stream)
{
    if (!E || !u || !dt) return ens_fail(E, PRANDTL_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    // bind the states and step sizes as a step would, then time the first
    // stage (it writes only the cases' stage buffers) in every launch form
    int rc = ens_bind(E, u, s);
    if (rc) return rc;
    if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
    if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
    DtArgs da;
    for (int i = 0; i < E->n; i++) da.dt[i] = dt[i];
    hipLaunchKernelGGL(k_set_dts, dim3(1), dim3(ENS_MAX_CASES), 0, s, E->d_dts, E->n, da);
    E->dt.assign(dt, dt + E->n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = ens_hip(E, hipEventCreate(&e0), "event");
    if (!rc) rc = ens_hip(E, hipEventCreate(&e1), "event");
    int best = 0;
    double best_us = 0.0;
    for (int mode = 0; mode <= PRANDTL_ENS_MAX_PER_CU && !rc; mode++) {
        constexpr int warm = 2, reps = 10;
        auto once = [&]() {
            if (mode > 0) hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, E->d_counters);
            ens_stage(E, 0, mode, s);
        };
        for (int i = 0; i < warm; i++) once();
        rc = ens_hip(E, hipEventRecord(e0, s), "event record");
        for (int i = 0; i < reps && !rc; i++) once();
        if (!rc) rc = ens_hip(E, hipGetLastError(), "k_ens launch");
        if (!rc) rc = ens_hip(E, hipEventRecord(e1, s), "event record");
        if (!rc) rc = ens_hip(E, hipEventSynchronize(e1), "event sync");
        float ms = 0.0f;
        if (!rc) rc = ens_hip(E, hipEventElapsedTime(&ms, e0, e1), "event time");
        if (rc) break;
        const double us = 1e3 * ms / reps;
        if (us_out) us_out[mode] = us;
        if (mode == 0 || us < best_us) { best = mode; best_us = us; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    E->mode = best;
    if (mode_out) *mode_out = best;
    return PRANDTL_OK;
}

const char *prandtl_ensemble_last_error(const prandtl_ensemble *E) { return E ? E->err.c_str() : g_create_err.c_str(); }

} // extern "C"
