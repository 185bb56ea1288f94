/*
 * prandtl_hip.h — C ABI of the MI355X-native DGSEM right-hand side and
 * SSP-RK3 stage update (library: prandtl_amd/libprandtl_hip.so).
 *
 * Drop-in boundary for Prandtl's semi-discrete operator on one GPU.  Every
 * entry point returns an int status (PRANDTL_OK = 0); no exception crosses
 * the ABI.  State vectors are device pointers in the reference's byNODES
 * layout: index = var*NE*K + elem*K + node, nodes lexicographic, x fastest
 * (Simulation.cpp:338-346 builds the vector space with Ordering::byNODES).
 * `stream` is a hipStream_t (NULL = default stream); calls are asynchronous
 * on it unless stated otherwise.  One context per stream/thread; a context
 * owns its device scratch, the caller owns the state buffers.
 */
#ifndef PRANDTL_HIP_H
#define PRANDTL_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PRANDTL_OK = 0,
    PRANDTL_ERR_ARG = 1,          /* bad argument or size */
    PRANDTL_ERR_UNSUPPORTED = 2,  /* configuration not built */
    PRANDTL_ERR_HIP = 3,          /* HIP runtime error (see prandtl_last_error) */
    PRANDTL_ERR_STATE = 4,        /* non-finite or non-positive density/pressure */
    PRANDTL_ERR_NOMEM = 5,
    PRANDTL_ERR_IO = 6            /* checkpoint file missing, truncated, corrupt or for another case */
};

enum { PRANDTL_FP64 = 0, PRANDTL_MIXED = 1 };

typedef struct prandtl_ctx prandtl_ctx;

/* Cartesian box of hexahedra/quadrilaterals (or a segment of intervals)
 * with LGL nodes.  Replaces the mesh/space set-up the operator receives in
 * DGSEMOperator::DGSEMOperator (DGSEMOperator.hpp:87-99; spaces built at
 * Simulation.cpp:338-346).  Element index = (ez*ny + ey)*nx + ex.  1-D
 * (the reference's TestCases/Euler/1D, MFEM inline segment meshes): Euler
 * only; face normals are +-1 and J = h (DGSEMIntegrator.cpp:180-183). */
typedef struct prandtl_mesh_desc {
    int dim;            /* 1, 2 or 3 */
    int order;          /* polynomial degree N (Np = N+1): 1..7 in 1-D, 1..5 in 2-D, 2..7 in 3-D */
    int nelem[3];       /* elements per direction */
    double lo[3], hi[3];
    /* Boundary condition of each box side: bc[r][0] at lo[r], bc[r][1] at
     * hi[r], PRANDTL_BC_* (zero: periodic).  A direction is periodic on both
     * sides or on neither.  The boundary integrators the reference attaches
     * by boundary name (Simulation.cpp:490-735); Euler only. */
    int bc[3][2];
    /* conservative outer state (dim + 2 values) of supersonic-inflow and
     * specified-state sides: the constant-state constructors of
     * SupersonicInflowBdrFaceIntegrator.cpp:12-17 and
     * SpecifiedStateBdrFaceIntegrator.cpp:12-16; wall velocity and heat
     * flux of no-slip sides (NoSlipAdiabWallBdrFaceIntegrator.cpp:10-15).
     * Navier-Stokes boxes take supersonic inflow / outflow, specified state
     * and no-slip walls (the reference has no viscous or lifting flux for
     * slip and symmetry sides); Euler boxes every type but no-slip walls. */
    double bc_state[3][2][5];
} prandtl_mesh_desc;

/* Boundary types, by the reference's configuration names ("type" of
 * runTime.conditions.boundary_conditions, Simulation.cpp:518-735) */
enum {
    PRANDTL_BC_PERIODIC = 0,
    PRANDTL_BC_SLIP = 1,               /* "slip": SlipWallBdrFaceIntegrator.cpp:15-60 */
    PRANDTL_BC_SUPERSONIC_INFLOW = 2,  /* "supersonic-inflow", constant state */
    PRANDTL_BC_SUPERSONIC_OUTFLOW = 3, /* "supersonic-outflow": outer state = interior */
    PRANDTL_BC_SYMMETRY = 4,           /* "symmetry": mirrored normal momentum */
    PRANDTL_BC_SPECIFIED_STATE = 5,    /* "specified-state", constant state */
    PRANDTL_BC_NO_SLIP_ADIABATIC = 6,  /* "no-slip-adiabatic" (Navier-Stokes): wall velocity in
                                          bc_state[0..dim-1], normal heat flux in bc_state[3];
                                          NoSlipAdiabWallBdrFaceIntegrator.cpp */
    PRANDTL_BC_AXIS = 7                /* "axis" (quadrilateral meshes): the symmetry flux, and in
                                          axisymmetric mode the marker of the axis nodes
                                          (Simulation.cpp:520-531) */
};

/* Physics and scheme switches: PhysicsConstants (Physics.hpp:10-37), the
 * compile-time PARABOLIC / SUBCELL_FV_BLENDING flags (CMakeLists.txt:36-50)
 * and DGSEMOperator's alpha_max/alpha_min (DGSEMOperator.hpp:99). */
typedef struct prandtl_params {
    double gamma, Pr, mu;
    double alpha_max, alpha_min;
    int parabolic;      /* Navier-Stokes with BR1 lifting */
    int blending;       /* modal indicator + subcell FV blending */
    int ec_surface;     /* 1: entropy-conservative two-point flux at faces */
    int quirks;         /* 1: reproduce the reference's defects, 0: corrected */
    int precision;      /* PRANDTL_FP64, or PRANDTL_MIXED: BR1 gradients, viscous
                           fluxes and their derivative sums in FP32 (entropy
                           variables, two-point fluxes, indicator and RK
                           accumulation stay FP64); Navier-Stokes with quirks = 0
                           only, identical to FP64 for Euler */
    int sutherland;     /* SUTHERLAND build (NavierStokesFlux.hpp:32-36): mu(T) by
                           Sutherland's law with mu0 = 1.716e-5, T0 = 273.15,
                           Ts = 110.4 (Physics.hpp:35-37), T = p / (rho R_gas);
                           mu is then unused */
    double R_gas;       /* gas constant for that temperature (> 0 with sutherland) */
    int axisymmetric;   /* AXISYMMETRIC build (quadrilateral meshes, Euler): the state
                           is r U with r = y; U is recovered from it before each RHS
                           (DGSEMOperator.cpp:161-424, with an "axis" boundary's
                           low-order axis reconstruction), fluxes are weighted by r
                           (DGSEMIntegrator.cpp:190-205,290,337,1030,1110,
                           BdrFaceIntegrator.cpp:109), the radial momentum gains the
                           pressure source (DGSEMIntegrator.cpp:411-426) and its RHS is
                           zeroed on the axis (DGSEMOperator.cpp:521-533,573-583) */
    double axis_rho_floor; /* density floor of the axis reconstruction (rho_floor_abs,
                           DGSEMOperator.hpp:44,134-138); <= 0: 1e-12 */
} prandtl_params;

/* DGSEMOperator constructor (DGSEMOperator.hpp:87-99): builds LGL rule,
 * derivative matrices (DGSEMIntegrator.cpp:8-82), modal basis
 * (ModalBasis.cpp:5-22) and the device scratch.  Uses the current device. */
int prandtl_create(const prandtl_mesh_desc *mesh, const prandtl_params *params,
                   prandtl_ctx **out);
void prandtl_destroy(prandtl_ctx *ctx);

/* ---- Unstructured 2-D quadrilateral meshes ----
 * The mesh the reference reads with mfem::Mesh(mesh_file) (Simulation.cpp:
 * 271-299) and its geometry: per-node Jacobians and adjugates of the
 * bilinear element map (used at DGSEMIntegrator.cpp:255-267), face normals
 * CalcOrtho(face Jacobian) outward of the face's first element
 * (DGSEMIntegrator.cpp:186, BdrFaceIntegrator.cpp:94), and the two
 * elements' face-point correspondence.  Elements are straight-sided quads
 * with vertices counterclockwise; element e's nodes are lexicographic in its
 * own frame (xi from vertex 0 to 1, eta from vertex 0 to 3). */
typedef struct prandtl_quad_mesh {
    int nvert;
    const double *vert;         /* [nvert][2] vertex coordinates */
    int nelem;
    const int *elem;            /* [nelem][4] vertex indices, counterclockwise */
    const double *elem_coords;  /* NULL, or [nelem][4][2] corner coordinates per
                                   element (periodic meshes: MFEM's L2 mesh nodes;
                                   topology then comes from elem) */
    int nbdr;
    const int *bdr;             /* [nbdr][2] boundary segments (vertex indices) */
    const int *bdr_attr;        /* [nbdr] boundary attribute, 1..nattr */
    int nattr;
    const int *attr_bc;         /* [nattr] PRANDTL_BC_* of each attribute */
    const double *attr_state;   /* [nattr][4] conservative outer states of the
                                   inflow / specified-state attributes (may be NULL
                                   if none uses one) */
} prandtl_quad_mesh;

/* DGSEMOperator constructor on a quadrilateral mesh (orders 1..5; Euler or
 * Navier-Stokes, with or without blending; axisymmetric Euler with
 * params->axisymmetric).  Every element edge must be shared by two elements
 * or be a boundary segment whose attribute has a non-periodic type.  Uses
 * the current device. */
int prandtl_create_quad(const prandtl_quad_mesh *mesh, int order, const prandtl_params *params,
                        prandtl_ctx **out);

/* ---- Hexahedral meshes ----
 * The 3-D counterpart (MFEM and Gmsh hexahedral meshes): trilinear
 * elements with vertices in MFEM's order (the zeta = 0 face counterclockwise
 * seen from +zeta, then the zeta = 1 face); per-node adjugates; faces paired
 * through their four vertices in any of the 8 relative orientations; the
 * normal at each face point is CalcOrtho of the face Jacobian (the first
 * element's adjugate row, outward).  Element e's nodes are lexicographic in
 * its own frame. */
typedef struct prandtl_hex_mesh {
    int nvert;
    const double *vert;         /* [nvert][3] */
    int nelem;
    const int *elem;            /* [nelem][8] vertex indices */
    const double *elem_coords;  /* NULL, or [nelem][8][3] (periodic meshes) */
    int nbdr;
    const int *bdr;             /* [nbdr][4] boundary quadrilaterals (vertex indices) */
    const int *bdr_attr;        /* [nbdr] boundary attribute, 1..nattr */
    int nattr;
    const int *attr_bc;         /* [nattr] PRANDTL_BC_* of each attribute */
    const double *attr_state;   /* [nattr][5] outer states (inflow / specified state:
                                   conservative; no-slip walls: wall velocity in
                                   [0..2], heat flux in [3]) */
} prandtl_hex_mesh;

/* DGSEMOperator constructor on a hexahedral mesh (orders 2..4; Euler or
 * Navier-Stokes, corrected or faithful, with or without blending).  Uses the
 * current device. */
int prandtl_create_hex(const prandtl_hex_mesh *mesh, int order, const prandtl_params *params,
                       prandtl_ctx **out);

/* Mesh files: mfem::Mesh's readers for "MFEM mesh v1.0" (vertices, or L2_T1
 * first-order nodes for periodic meshes) and Gmsh 2.2 ASCII ($PhysicalNames,
 * $Nodes, $Elements with 4-node quadrilaterals and 2-node boundary lines, or
 * 8-node hexahedra and 4-node boundary quadrilaterals; attributes are the
 * physical tags; a $Periodic section identifies slave nodes with their
 * masters and keeps per-element corner coordinates, as mfem::Mesh does).  Elements keep file order.  Elements
 * with clockwise vertices are reoriented.  _refine splits every element into
 * four (children of element i at 4i..4i+3, corners in the parent's frame
 * (0,0), (1,0), (1,1), (0,1)), a hexahedron into eight (children at
 * 8i..8i+7, child (a,b,c) at 8i + a + 2b + 4c), as ser_ref_levels does with
 * Mesh::UniformRefinement (Simulation.cpp:293-300).  _view fills the
 * geometry fields of a prandtl_quad_mesh (pointers into the file object;
 * the caller sets nattr / attr_bc / attr_state).  _name returns the name of
 * physical group i (0-based) and its attribute, NULL past the end. */
typedef struct prandtl_mesh_file prandtl_mesh_file;
int prandtl_mesh_read(const char *path, prandtl_mesh_file **out);
int prandtl_mesh_refine(prandtl_mesh_file *m);
int prandtl_mesh_view(const prandtl_mesh_file *m, prandtl_quad_mesh *view);
/* 2 for a quadrilateral mesh, 3 for a hexahedral one (Gmsh 8-node
 * hexahedra with 4-node boundary quadrilaterals and 2-D physical names, or
 * MFEM geometry 5 with boundary geometry 3); _view_hex fills a
 * prandtl_hex_mesh (periodic ones come from Gmsh's $Periodic section or
 * MFEM's L2_T1_3D_P1 node sets) */
int prandtl_mesh_dim(const prandtl_mesh_file *m);
int prandtl_mesh_view_hex(const prandtl_mesh_file *m, prandtl_hex_mesh *view);
const char *prandtl_mesh_name(const prandtl_mesh_file *m, int i, int *attr);
void prandtl_mesh_free(prandtl_mesh_file *m);
const char *prandtl_mesh_last_error(void);

/* NE, K (nodes/element), nv (variables) */
int prandtl_sizes(const prandtl_ctx *ctx, int *NE, int *K, int *nv);

/* the smallest element size h_min of the dt policy: min over elements of
 * GetElementSize(e, 1), the Jacobian's smaller singular value at the element
 * centre (Simulation.cpp:896-901); min_r (hi-lo)/nelem on a box */
int prandtl_hmin(const prandtl_ctx *ctx, double *hmin);

/* DGSEMOperator::Mult(const Vector &u, Vector &dudt) const
 * (DGSEMOperator.hpp:105, DGSEMOperator.cpp:537-591): dudt is overwritten
 * with the semi-discrete RHS of u.  u and dudt must not alias. */
int prandtl_rhs(prandtl_ctx *ctx, const double *u_dev, double *dudt_dev, void *stream);

/* No counterpart in the reference: the RHS of prandtl_rhs on a Cartesian box
 * (periodic or with boundary sides) for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It takes no context: the operator's tables are built on
 * the host at each call and passed to the kernels by value, and the kernels'
 * scratch (k_visc's viscous sums and face traces, the wave-speed slots,
 * alpha, eta, the box's boundary states) lives in a fixed 32 MiB device array
 * of the library, one per GPU, held until the process ends.  The call makes
 * no hipMalloc, hipFree or other allocating call on the device; it returns
 * PRANDTL_ERR_ARG for a case whose scratch exceeds 32 MiB (3-D
 * Navier-Stokes: about 400k nodes; the message gives the bytes needed) and
 * the create-time errors of prandtl_create.  The result equals prandtl_rhs
 * of a context created from the same mesh and params, bitwise.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other host threads) race and give wrong results.  The running
 * wave speed is not kept.  Errors: prandtl_last_error(NULL). */
int prandtl_rhs_noalloc(const prandtl_mesh_desc *mesh, const prandtl_params *params, const double *u_dev,
                        double *dudt_dev, void *stream);

/* RK3SSPSolver::Step(x, t, dt) as driven at Simulation.cpp:1010: three RHS
 * evaluations with the Shu-Osher stage combinations fused into the RHS
 * kernel; u is advanced in place. */
int prandtl_rk3_step(prandtl_ctx *ctx, double *u_dev, double dt, void *stream);

/* Simulation::Run's stepping (Simulation.cpp:893-907, 1007-1022) on the
 * device: SSP-RK3 steps of u_dev from *t_io until t_final, each step
 * clipped to t_final - t.  cfl > 0: variable dt = cfl*hmin/(lambda*(N+1)^2)
 * from the running maximum wave speed, set after an initial RHS evaluation
 * and after every step; otherwise the fixed step dt.  Stops when
 * t >= t_final - 1e-8 dt, or after max_steps steps if max_steps > 0.  No host
 * synchronisation per step: batches of 16 steps run from one graph (on a
 * non-null stream) and the host reads the loop state once per batch; steps
 * after the end do nothing.  *t_io receives the final time, *steps_out
 * (may be NULL) the number of steps.  Bitwise the same steps as the host
 * loop of prandtl_rk3_step and prandtl_max_wave_speed.  check_every > 0:
 * after every step whose count is a multiple of it, the NaN/Inf check of
 * Simulation.cpp:1025-1035 (extended to rho <= 0 and p <= 0) stops the loop
 * with PRANDTL_ERR_STATE; *t_io and *steps_out then give where it stopped. */
int prandtl_run(prandtl_ctx *ctx, double *u_dev, double *t_io, double t_final, double cfl,
                double dt, int max_steps, int check_every, int *steps_out, void *stream);

/* On a non-null stream the three stages of prandtl_rk3_step are captured
 * once as a hipGraph (keyed by u_dev and stream; dt is read from device
 * memory) and replayed on later steps.  enable = 0 launches the kernels
 * directly every step.  Default: enabled. */
int prandtl_set_graph(prandtl_ctx *ctx, int enable);

/* Launch shapes (no counterpart in the reference): how many elements one
 * workgroup of a kernel takes and how many threads it has.  kernel is
 * PRANDTL_KERNEL_RHS (the flux-differencing kernel) or PRANDTL_KERNEL_VISC
 * (the BR1 lifting and viscous-flux kernel, Navier-Stokes only).  Shape 0 is
 * the default; shapes 1..3 force 64, 128 or 256 threads per workgroup where
 * that still gives every node its own thread.  Every shape gives bitwise the
 * same results.  PRANDTL_ERR_UNSUPPORTED for a shape this case does not
 * have. */
#define PRANDTL_SHAPES 4
#define PRANDTL_KERNEL_RHS 0
#define PRANDTL_KERNEL_VISC 1
int prandtl_set_shape(prandtl_ctx *ctx, int kernel, int shape);

/* Kernel algorithm of the flux-differencing RHS (no counterpart in the
 * reference; the operator is the same, DGSEMIntegrator.cpp:151-1196).
 * PRANDTL_ALGO_TILE: k_rhs, a workgroup's threads take the element's
 * two-point fluxes as jobs through LDS (every mesh kind).  PRANDTL_ALGO_LINE:
 * k_rhs_ln, one thread per tensor line (periodic Cartesian boxes in 3-D,
 * N = 2..4), the default at N = 2 and 3 and, with blending, at N = 4 (the
 * faster of the two there; unblended N = 4 defaults to the tile kernel);
 * with it, the mixed form at N = 3
 * runs its viscous kernel's derivative contractions on the FP32 matrix cores
 * (k_visc_mx).  The two differ by rounding only;
 * launch shapes (prandtl_set_shape, prandtl_autotune) belong to the current
 * algorithm (the line kernel has shape 0 only).  PRANDTL_ERR_UNSUPPORTED for
 * an algorithm this case does not have. */
#define PRANDTL_ALGO_TILE 0
#define PRANDTL_ALGO_LINE 1
int prandtl_set_algorithm(prandtl_ctx *ctx, int kernel, int algo);
int prandtl_get_algorithm(const prandtl_ctx *ctx, int kernel, int *algo);

/* Autotuner: times every shape this case has, of each kernel, on state
 * u_dev (2 + 10 launches each, into context scratch; the running wave speed
 * sees u_dev) and keeps the fastest of each.  shapes_out (may be NULL)
 * receives the two chosen shapes [rhs, visc]; us_out (may be NULL) receives
 * 2 * PRANDTL_SHAPES mean microseconds per launch, rhs shapes first, -1 for
 * shapes a case does not have. */
int prandtl_autotune(prandtl_ctx *ctx, const double *u_dev, int *shapes_out, double *us_out,
                     void *stream);

/* TimeDependentOperator::SetTime / GetTime (DGSEMOperator::GetTimeRef,
 * DGSEMOperator.hpp:111-114): the time of the next RHS.  prandtl_rhs
 * evaluates at it; prandtl_rk3_step evaluates its stages at t, t + dt and
 * t + dt/2 (RK3SSPSolver's SetTime calls, Simulation.cpp:1010) and advances
 * it by dt; prandtl_run runs from and leaves it at the loop's time.  Only
 * time-dependent boundary states read it.  set_time orders the update on
 * `stream` (a hipStream_t). */
int prandtl_set_time(prandtl_ctx *ctx, double t, void *stream);
int prandtl_get_time(const prandtl_ctx *ctx, double *t);

/* A time-dependent outer state of boundary attribute attr (1-based) of a 2-D
 * quadrilateral mesh, for its specified-state or supersonic-inflow type: the
 * reference's vector TD-function boundary conditions (Simulation.cpp:
 * 541-563, "time_dependent": true; SpecifiedStateBdrFaceIntegrator.cpp:
 * 24-34 evaluates the function at each boundary point and the operator's
 * time), in the form they take in the reference's cases: a straight
 * discontinuity moving at constant speed.  At a boundary point x and stage
 * time t the outer conservative state is below[] where
 * normal[0] x0 + normal[1] x1 - offset - speed t < 0 and above[] elsewhere
 * (DoubleMachReflectionTopBCFunction, TestCases/Euler/2D/
 * DoubleMachReflection/DoubleMachReflection.hpp:33-57: normal = (1,
 * -1/sqrt(3)), offset 1/6, speed 20/sqrt(3)).  Euler, non-axisymmetric
 * contexts; PRANDTL_ERR_UNSUPPORTED otherwise. */
int prandtl_set_boundary_split(prandtl_ctx *ctx, int attr, const double *normal, double offset, double speed,
                               const double *below, const double *above);

/* DGSEMOperator::GetMaxCharSpeed (DGSEMOperator.hpp:106-109): running
 * maximum of |v|+c over every state the operator has evaluated since
 * creation or the last reset (DGSEMIntegrator.cpp:39,188,270).  Synchronises
 * the stream.  reset != 0 clears it after reading. */
int prandtl_max_wave_speed(prandtl_ctx *ctx, double *out, int reset, void *stream);

/* Blending coefficient alpha_e and indicator eta_e of the last RHS
 * (the `alpha` / `eta` grid functions of DGSEMOperator.cpp:60-80), copied to
 * host arrays of NE doubles.  Synchronises the stream. */
int prandtl_alpha(prandtl_ctx *ctx, double *alpha_host, double *eta_host, void *stream);

/* Simulation.cpp:1025-1035 nancheck, extended: returns PRANDTL_ERR_STATE if
 * any node has a non-finite value, rho <= 0 or p <= 0.  Synchronises. */
int prandtl_check_state(prandtl_ctx *ctx, const double *u_dev, void *stream);

/* LGL nodes and weights on [0,1] used by the operator (DGSEMIntegrator.cpp:18-22). */
int prandtl_lgl(int order, double *x, double *w);

/* Diagnostics of the parabolic path: primitive-variable gradients
 * [dim][nv][NE][K] of u after BR1 lifting and EntropyGrad2PrimGrad
 * (DGSEMOperator.cpp:96-159), written to a device buffer. */
int prandtl_debug_gradients(prandtl_ctx *ctx, const double *u_dev, double *grad_dev, void *stream);

/* Kernel timing, the per-phase timing the reference lacks (it has only a
 * whole-run clock, Simulation.cpp:910-914,1162-1171): between begin and end
 * every kernel launch of this context is bracketed by HIP events on its
 * stream.  end synchronises and returns the summed durations in ms and the
 * launch counts: [0] = RHS/stage kernel, [1] = parabolic (lifting/viscous)
 * kernel. */
int prandtl_profile_begin(prandtl_ctx *ctx);
int prandtl_profile_end(prandtl_ctx *ctx, double *ms, int *count);

/* Achievable machine peaks of the current device (no counterpart in the
 * reference), for the roofline beside the datasheet figures.
 * prandtl_peak_fma: vector FMA rate in TFLOP/s of bits = 64 (FP64) or 32
 * (unpacked FP32), from 8 independent FMA chains per thread at 8 waves per
 * SIMD; sink is a device buffer of at least 8 bytes that is never written.
 * prandtl_peak_copy: mean microseconds of reps device-to-device copies of
 * bytes (a multiple of 16) from src to dst, device buffers that do not
 * overlap (read and write traffic: 2 bytes per byte copied). */
int prandtl_peak_fma(int bits, void *sink_dev, double *tflops, void *stream);
int prandtl_peak_copy(const void *src_dev, void *dst_dev, size_t bytes, int reps, double *us, void *stream);

/* Last error message of this context (or of create when ctx is NULL). */
const char *prandtl_last_error(const prandtl_ctx *ctx);

/* Primitive output fields (Simulation.cpp:1046-1061): prim_dev receives
 * rho, the velocity components and p, [dim+2][NE][K] like the state. */
int prandtl_primitives(prandtl_ctx *ctx, const double *u_dev, double *prim_dev, void *stream);

/* DGSEMOperator::RecoverStateFromWeighted (DGSEMOperator.cpp:161-424): the
 * conservative state U_dev of the weighted state rU_dev of an axisymmetric
 * context (both [nv][NE][K] device buffers); prandtl_primitives of such a
 * context applies it first, as the reference's output does
 * (Simulation.cpp:1041-1044). */
int prandtl_axis_recover(prandtl_ctx *ctx, const double *rU_dev, double *U_dev, void *stream);

/* ---- Checkpoints (Simulation.cpp:1084-1116 save, 350-388 load) ----
 * The folder layout of the reference: dir/Cycle<cycle>/checkpoint_cycle_<cycle>.json
 * holding {"time", "cycle"}, and .../checkpoint_cycle_<cycle>.00000000.chk.
 * The .chk is this library's format, not MFEM's GridFunction text: a header
 * (magic, version, dim, order, elements per direction, variables, value
 * count), the state as raw FP64 in byNODES order, and a 64-bit FNV-1a
 * checksum of the state bytes.  The time is written at full precision.
 * Reading checks the header against the expected mesh and variables, the
 * length and the checksum; any mismatch is PRANDTL_ERR_IO with the reason in
 * prandtl_last_error.  _write/_read work on host arrays and need no GPU;
 * _save/_load copy a context's device state (u_dev) through them. */
int prandtl_checkpoint_write(const char *dir, int cycle, double t, const prandtl_mesh_desc *mesh, int nv,
                             const double *u_host);
int prandtl_checkpoint_read(const char *dir, int cycle, const prandtl_mesh_desc *mesh, int nv,
                            double *u_host, double *t_out);
int prandtl_checkpoint_save(prandtl_ctx *ctx, const double *u_dev, const char *dir, int cycle, double t,
                            void *stream);
int prandtl_checkpoint_load(prandtl_ctx *ctx, double *u_dev, const char *dir, int cycle, double *t_out,
                            void *stream);

/* ---- Batched ensemble (no counterpart in the reference, which runs one
 * case per process: Simulation::Run, Simulation.cpp:882-1181) ----
 * Independent 2-D Euler cases of orders 2..5, with or without blending, each
 * with its own mesh and parameters.  One RK3 step advances every case: per
 * stage a single launch covers the element tiles of all cases, one tile per
 * workgroup, in a host-built order of decreasing cost that the workgroup
 * dispatcher consumes as a work queue; on a non-null stream the three stages
 * are captured once as a hipGraph.  Results are identical to stepping each
 * case alone with prandtl_rk3_step.  At most 256 cases.  The case contexts
 * belong to the ensemble; prandtl_ensemble_destroy frees them. */
typedef struct prandtl_ensemble prandtl_ensemble;
int prandtl_ensemble_create(int ncases, const prandtl_mesh_desc *meshes,
                            const prandtl_params *params, prandtl_ensemble **out);
void prandtl_ensemble_destroy(prandtl_ensemble *ens);
/* the per-case context (sizes, wave speed, alpha) */
int prandtl_ensemble_case(prandtl_ensemble *ens, int i, prandtl_ctx **ctx);
/* u: ncases device state pointers, dt: ncases step sizes (host array) */
int prandtl_ensemble_rk3_step(prandtl_ensemble *ens, double *const *u_dev, const double *dt,
                              void *stream);
const char *prandtl_ensemble_last_error(const prandtl_ensemble *ens);

/* Launch form of the ensemble's stages (no counterpart in the reference):
 * mode 0 launches one workgroup per element tile (the cost-sorted tile list
 * is the dispatch order); mode g = 1..PRANDTL_ENS_MAX_PER_CU launches g
 * persistent workgroups per CU that take tiles from a device counter until
 * the list is exhausted.  Every form gives bitwise the same results. */
#define PRANDTL_ENS_MAX_PER_CU 4
int prandtl_ensemble_set_mode(prandtl_ensemble *ens, int mode);
/* Autotuner of the launch form: binds u_dev / dt as prandtl_ensemble_rk3_step
 * would and times the first RK stage (which writes only the cases' stage
 * buffers, not u_dev) in every form, 2 + 10 launches each; keeps the fastest.
 * mode_out (may be NULL) receives it, us_out (may be NULL) the
 * PRANDTL_ENS_MAX_PER_CU + 1 mean microseconds per stage, form 0 first. */
int prandtl_ensemble_autotune(prandtl_ensemble *ens, double *const *u_dev, const double *dt, int *mode_out,
                              double *us_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
