/*
 * TEST INFRASTRUCTURE — not part of the product.
 *
 * C ABI of the CPU FP64 restatement ("the oracle") of Prandtl's DGSEM
 * right-hand side and SSP-RK3 step.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library, and only as the
 * checker / CPU baseline; the product path never links or calls it.
 *
 * Layout of every state vector: byNODES SoA, index = var*NE*K + e*K + node,
 * nodes lexicographic with x fastest (Simulation.cpp:338-346 in the
 * reference; vdim = nv, Ordering::byNODES).
 */
#ifndef PRANDTL_ORACLE_H
#define PRANDTL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

typedef struct orc_params {
    int dim;            /* 2 or 3 */
    int order;          /* polynomial degree N; Np = N+1 LGL nodes per direction */
    int nelem[3];       /* periodic Cartesian mesh: elements per direction */
    double lo[3], hi[3];/* box */
    double gamma, Pr, mu;
    int parabolic;      /* Navier-Stokes with BR1 lifting (PARABOLIC build) */
    int blending;       /* SUBCELL_FV_BLENDING build */
    double alpha_max, alpha_min;
    int ec_surface;     /* 1: entropy-conservative F# at faces (entropy test) */
    int quirks;         /* 1: reproduce the reference's defects, 0: corrected */
    int threads;        /* OpenMP threads for the double instantiation (<=0: 1) */
    int sutherland;     /* SUTHERLAND build: mu(T) by Sutherland's law, T = p / (rho R_gas) */
    double R_gas;       /* gas constant of that temperature (> 0 with sutherland) */
    /* boundary condition of each box side, bc[r][0] at lo[r], bc[r][1] at
     * hi[r]: ORC_BC_* below; a direction is periodic on both sides or on
     * neither.  bc_state: the constant conservative outer state of the
     * supersonic-inflow / specified-state sides.  Euler only. */
    int bc[3][2];
    double bc_state[3][2][5];
    /* AXISYMMETRIC build (quadrilateral meshes, Euler): the state is r U with
     * r = y, U is recovered from it before each RHS (DGSEMOperator.cpp:
     * 161-424), fluxes are weighted by r and the radial momentum gains the
     * pressure source.  axis_rho_floor: the density floor of the axis
     * reconstruction (rho_floor_abs, DGSEMOperator.hpp:44; <= 0: 1e-12) */
    int axisymmetric;
    double axis_rho_floor;
    /* TEST-ONLY: a bitmask of ORC_PLANT_* misreadings of the faithful-mode
     * branches and of the indicator.  The pins of tests/test_oracle_faithful.py
     * plant each one and show that they fail on it; every other use leaves
     * it 0. */
    int plant;
} orc_params;

/* planted misreadings (orc_params.plant) */
enum {
    ORC_PLANT_STALE_NODE0 = 1,     /* A-3: the viscous volume flux takes node 0's state, not node K-1's */
    ORC_PLANT_ZETA_F = 2,          /* A-2: the zeta partner's column takes its own flux h, not g */
    ORC_PLANT_ZETA_XI = 4,         /* A-2: g is the last xi-direction flux, not the last eta one */
    ORC_PLANT_FACE_NO_ALIAS = 8,   /* A-8: the Navier-Stokes face flux reads momenta, not velocities */
    ORC_PLANT_GRAD_NO_ALIAS = 16,  /* A-8: EntropyGrad2PrimGrad's density row multiplies momenta */
    ORC_PLANT_THRESH_NP1 = 32      /* indicator threshold 0.5 10^(-1.8 (N+1)^0.25) instead of N */
};

enum {
    ORC_BC_PERIODIC = 0,
    ORC_BC_SLIP = 1,               /* SlipWallBdrFaceIntegrator.cpp:15-60 */
    ORC_BC_SUPERSONIC_INFLOW = 2,  /* SupersonicInflowBdrFaceIntegrator.cpp (constant state) */
    ORC_BC_SUPERSONIC_OUTFLOW = 3, /* SupersonicOutflowBdrFaceIntegrator.cpp */
    ORC_BC_SYMMETRY = 4,           /* SymmetryBdrFaceIntegrator.cpp */
    ORC_BC_SPECIFIED_STATE = 5,    /* SpecifiedStateBdrFaceIntegrator.cpp (constant state) */
    ORC_BC_NO_SLIP_ADIABATIC = 6,  /* NoSlipAdiabWallBdrFaceIntegrator.cpp (constant wall velocity
                                      bc_state[0..dim-1] and heat flux bc_state[3]) */
    ORC_BC_AXIS = 7                /* "axis": the symmetry flux, and in axisymmetric mode the
                                      marker of the axis nodes (Simulation.cpp:520-531) */
};

/* 1-D rules and operators (setup arithmetic of DGSEMIntegrator.cpp:8-82) */
int orc_lgl(int order, double *x, double *w);
int orc_dmat(int order, double *D, double *Dhat, double *Dhat2); /* row-major Np x Np */
int orc_legendre_vdm(int order, double *V1, double *V1inv);      /* row-major 1-D shifted-Legendre Vandermonde */

/* point physics (ChandrashekarFlux.cpp, BasicOperations.cpp, NavierStokesFlux.cpp) */
double orc_logmean(double x, double y);
double orc_volume_flux(int dim, double gamma, const double *s1, const double *s2,
                       const double *m1, const double *m2, double *F);
double orc_face_flux(int dim, double gamma, int quirks, const double *s1, const double *s2,
                     const double *nor, double *F);
void orc_cons2entropy(int dim, double gamma, const double *s, double *w);
void orc_entgrad2primgrad(int dim, double gamma, int quirks, const double *s, double *g);
/* R_gas > 0: Sutherland viscosity (mu is then unused) */
void orc_viscous_flux(int dim, double gamma, double Pr, double mu, double R_gas, const double *s,
                      const double *dqdx, const double *dqdy, const double *dqdz,
                      double *flux /* nv x dim row-major */);

/* ComputeBdrFaceInviscidFlux of boundary type `type` for the interior state
 * s1, outer state `outer` (inflow / specified state only) and the outward
 * face normal nor (face-measure scaled); returns the wave speed */
double orc_bc_flux(int dim, double gamma, int quirks, int type, const double *s1, const double *outer,
                   const double *nor, double *F);
/* ComputeBdrFaceLiftingFlux of boundary type `type` for the interior entropy
 * variables w (BdrFaceIntegrator.cpp:276-282, NoSlipAdiabWallBdrFaceIntegrator.cpp) */
void orc_bc_lift_flux(int dim, double gamma, int quirks, int type, const double *w, const double *outer,
                      double *f);

/* operator */
int orc_create(const orc_params *p, orc_ctx **out);

/* 2-D quadrilateral mesh: straight-sided elements, vertices counterclockwise,
 * element e's nodes lexicographic in its own frame (xi from vertex 0 to 1,
 * eta from vertex 0 to 3).  elem_coords (optional) gives per-element corner
 * coordinates for periodic meshes, topology then from elem.  Boundary
 * segments carry attributes 1..nattr, whose ORC_BC_* type and outer state
 * (4 values) are attr_bc / attr_state. */
typedef struct orc_quad_mesh {
    int nvert;
    const double *vert;
    int nelem;
    const int *elem;
    const double *elem_coords;
    int nbdr;
    const int *bdr;
    const int *bdr_attr;
    int nattr;
    const int *attr_bc;
    const double *attr_state;
} orc_quad_mesh;
/* p->dim must be 2; p->nelem / lo / hi / bc are ignored */
int orc_create_quad(const orc_params *p, const orc_quad_mesh *m, orc_ctx **out);

/* 3-D hexahedral mesh: trilinear elements, vertices in MFEM's order (the
 * zeta = 0 face counterclockwise seen from +zeta, then zeta = 1); element
 * e's nodes lexicographic in its own frame.  elem_coords (optional) gives
 * per-element corner coordinates [nelem][8][3] for periodic meshes.
 * Boundary quadrilaterals (4 vertices each) carry attributes 1..nattr;
 * attr_state holds 5 values per attribute.  p->dim must be 3. */
typedef struct orc_hex_mesh {
    int nvert;
    const double *vert;
    int nelem;
    const int *elem;
    const double *elem_coords;
    int nbdr;
    const int *bdr;
    const int *bdr_attr;
    int nattr;
    const int *attr_bc;
    const double *attr_state;
} orc_hex_mesh;
int orc_create_hex(const orc_params *p, const orc_hex_mesh *m, orc_ctx **out);
void orc_destroy(orc_ctx *c);
int orc_sizes(const orc_ctx *c, int *NE, int *K, int *nv);
int orc_coords(const orc_ctx *c, double *X /* [dim][NE][K] */);
int orc_weights(const orc_ctx *c, double *WJ /* [NE][K] omega_i * J_i */);
int orc_rhs(orc_ctx *c, const double *u, double *dudt);
/* the same RHS evaluated in x87 extended precision (64-bit mantissa), as a
 * higher-precision reference for rounding-error studies */
int orc_rhs_ld(orc_ctx *c, const double *u, double *dudt);
int orc_rk3_step(orc_ctx *c, double *u, double dt);
/* TimeDependentOperator::SetTime/GetTime: the time of the next RHS; an RK3
 * step evaluates its stages at t, t + dt, t + dt/2 (MFEM RK3SSPSolver) and
 * advances the time by dt */
int orc_set_time(orc_ctx *c, double t);
double orc_get_time(const orc_ctx *c);
/* A time-dependent outer state of boundary attribute attr (1-based) of a
 * quadrilateral mesh (specified-state or supersonic-inflow boundaries, the
 * reference's vector TD-function conditions, Simulation.cpp:541-563,
 * SpecifiedStateBdrFaceIntegrator.cpp:24-34): below[] at the boundary points
 * x with n . x - offset - speed t < 0, above[] elsewhere */
int orc_set_boundary_split(orc_ctx *c, int attr, const double *normal, double offset, double speed,
                           const double *below, const double *above);
double orc_max_wave_speed(const orc_ctx *c);
void orc_reset_wave_speed(orc_ctx *c);
int orc_alpha(const orc_ctx *c, double *alpha /* NE */);
int orc_eta(const orc_ctx *c, double *eta /* NE */);
double orc_hmin(const orc_ctx *c);
/* DGSEMOperator::RecoverStateFromWeighted (DGSEMOperator.cpp:161-424): the
 * conservative state U of the weighted state rU (axisymmetric contexts) */
int orc_axis_recover(const orc_ctx *c, const double *rU, double *U);
/* BR1 primitive gradients of the last parabolic RHS, [dim][nv][NE][K] */
int orc_last_gradients(const orc_ctx *c, double *grad);
/* algorithmic op count of one RHS on state u (counted instantiation):
 * out[0]=total, [1]=indicator+alpha, [2]=entropy vars+lifting+grad conversion,
 * [3]=viscous volume+face viscous fluxes, [4]=inviscid volume pairs,
 * [5]=surface fluxes, [6]=FV subcell, [7]=assembly (scaling/sums) */
int orc_count_rhs_flops(orc_ctx *c, const double *u, double *out);

#ifdef __cplusplus
}
#endif
#endif
