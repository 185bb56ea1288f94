"""Synthetic meshes and initial states of the benchmarked configurations.

All states are numpy float64 arrays in byNODES layout [var][element][node]
(element index (ez*ny + ey)*nx + ex, nodes lexicographic x fastest) on the
periodic Cartesian boxes the operator builds.

 - isentropic_vortex: 2-D Euler, Shu vortex of strength beta on [-5,5]^2
   around the freestream rho=1, u=v=1, p=1.
 - taylor_green: 3-D Taylor-Green vortex, rho=1, V0=1, L=1,
   p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16, p0 = 1/(gamma Ma^2).
 - kelvin_helmholtz: 2-D shear layer, B = tanh(s y + s/2) - tanh(s y - s/2),
   rho = 0.5 + 0.75 B, u = 0.5 (B - 1), p = 1,
   v = 0.1 sin(2 pi x) + 1e-3 (2 r - 1) with r from std::mt19937_64
   (seed 20261202) as (x >> 11) * 2^-53, drawn in global node order.
"""
import numpy as np

from .operator import lgl_nodes

MASK64 = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (the C++ standard's 64-bit Mersenne Twister)."""

    def __init__(self, seed=5489):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK64
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK64
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK64

    def uniform(self):
        """(x >> 11) * 2^-53 in [0, 1)."""
        return (self() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_array(self, n):
        """The next n values of uniform(), drawn 312 at a time with numpy.

        The twist splits into three runs whose inputs are all known when the
        run starts: i < 156 reads old words only, 156 <= i < 311 reads
        mt[i-156] (updated by the first run), and i = 311 reads the new mt[0].
        """
        out = np.empty(n, dtype=np.uint64)
        mt = np.array(self.mt, dtype=np.uint64)
        upper, lower = np.uint64(0xFFFFFFFF80000000), np.uint64(0x7FFFFFFF)
        matrix = np.uint64(0xB5026F5AA96619E9)
        one = np.uint64(1)

        def twist_run(i0, i1, nxt, far):
            x = (mt[i0:i1] & upper) | (nxt & lower)
            xa = (x >> one) ^ np.where((x & one) == one, matrix, np.uint64(0))
            mt[i0:i1] = far ^ xa

        k = 0
        idx = self.idx
        while k < n:
            if idx >= 312:
                twist_run(0, 156, mt[1:157], mt[156:312])
                twist_run(156, 311, mt[157:312], mt[0:155])
                twist_run(311, 312, mt[0:1], mt[155:156])
                idx = 0
            take = min(312 - idx, n - k)
            y = mt[idx:idx + take].copy()
            y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
            y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
            y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
            y ^= y >> np.uint64(43)
            out[k:k + take] = y
            k += take
            idx += take
        self.mt = [int(v) for v in mt]
        self.idx = idx
        return (out >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def node_coords(dim, order, nelem, lo, hi):
    """Physical node coordinates X[dim][NE][K]."""
    x, _ = lgl_nodes(order)
    Np = order + 1
    n = [int(nelem[r]) for r in range(dim)] + [1] * (3 - dim)
    h = [(hi[r] - lo[r]) / n[r] for r in range(dim)]
    NE = int(np.prod(n))
    K = Np ** dim
    e = np.arange(NE)
    eidx = [e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])]
    i = np.arange(K)
    iidx = [i % Np, (i // Np) % Np, i // (Np * Np)]
    X = np.zeros((dim, NE, K))
    for r in range(dim):
        X[r] = lo[r] + (eidx[r][:, None] + x[iidx[r]][None, :]) * h[r]
    return X


def prim_to_cons(rho, vel, p, gamma):
    dim = len(vel)
    U = np.zeros((dim + 2,) + np.shape(rho))
    U[0] = rho
    ke = np.zeros_like(rho)
    for d in range(dim):
        U[1 + d] = rho * vel[d]
        ke = ke + vel[d] * vel[d]
    U[dim + 1] = p / (gamma - 1.0) + 0.5 * rho * ke
    return U


def isentropic_vortex(X, gamma=1.4, beta=5.0, center=(0.0, 0.0), uinf=(1.0, 1.0)):
    x = X[0] - center[0]
    y = X[1] - center[1]
    r2 = x * x + y * y
    f = beta / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    u = uinf[0] - f * y
    v = uinf[1] + f * x
    T = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T
    return prim_to_cons(rho, [u, v], p, gamma)


def isentropic_vortex_exact(X, t, lo, hi, gamma=1.4, beta=5.0, uinf=(1.0, 1.0)):
    """Exact solution at time t: the vortex advected with uinf, periodically wrapped."""
    L = [hi[0] - lo[0], hi[1] - lo[1]]
    c = [uinf[0] * t, uinf[1] * t]
    x = X[0] - c[0]
    y = X[1] - c[1]
    x = (x - lo[0]) % L[0] + lo[0]
    y = (y - lo[1]) % L[1] + lo[1]
    return isentropic_vortex(np.stack([x, y]), gamma, beta, (0.0, 0.0), uinf)


def taylor_green(X, gamma=1.4, Ma=0.1):
    x, y, z = X[0], X[1], X[2]
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    p0 = 1.0 / (gamma * Ma * Ma)
    p = p0 + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    rho = np.ones_like(u)
    return prim_to_cons(rho, [u, v, w], p, gamma)


def kelvin_helmholtz(X, gamma=1.4, s=15.0, seed=20261202, noise=1e-3):
    x, y = X[0], X[1]
    B = np.tanh(s * y + 0.5 * s) - np.tanh(s * y - 0.5 * s)
    rho = 0.5 + 0.75 * B
    u = 0.5 * (B - 1.0)
    r = MT19937_64(seed).uniform_array(x.size).reshape(x.shape)
    v = 0.1 * np.sin(2.0 * np.pi * x) + noise * (2.0 * r - 1.0)
    p = np.ones_like(x)
    return prim_to_cons(rho, [u, v], p, gamma)


def triple_point(X, gamma=1.4):
    """TriplePointShockInteractionIC (TestCases/Euler/2D/TriplePointShockInteraction/
    TriplePointShockInteraction.hpp:9-39), gas at rest: rho = 1, p = 1 for
    x <= 1; rho = 1, p = 0.1 for x <= 7 and y <= 1.5; rho = 0.125, p = 0.1
    elsewhere."""
    x, y = X[0], X[1]
    left = x <= 1.0
    low = (x <= 7.0) & (y <= 1.5)
    rho = np.where(left, 1.0, np.where(low, 1.0, 0.125))
    p = np.where(left, 1.0, 0.1)
    zero = np.zeros_like(x)
    return prim_to_cons(rho, [zero, zero], p, gamma)


# The triple-point problem of the reference's test cases (config.json there:
# order 3, SUBCELL_FV_BLENDING, slip walls on all four sides, cfl 1.0, final
# time 4.0) on the box [0,7] x [0,3]; its Gmsh mesh is not shipped, so the box
# is meshed uniformly (h = 1/16 here).
TRIPLE_POINT = dict(dim=2, order=3, nelem=(112, 48), lo=(0.0, 0.0), hi=(7.0, 3.0), gamma=1.4,
                    blending=True, alpha_max=0.5, bc=[("slip", "slip"), ("slip", "slip")])


# The reference's 1-D Euler cases (TestCases/Euler/1D: order 3,
# SUBCELL_FV_BLENDING, cfl 1.0, an MFEM inline segment mesh of nx elements on
# [0, sx] refined ser_ref_levels times).  Each is a piecewise state: `states`
# holds (x_end, rho, u, p) pieces in order, x <= x_end for "le" cases, x <
# x_end otherwise (rho may be a function of x); specified-state and inflow
# sides take the state the initial condition gives at that end.  gamma and
# alpha_max default to 1.4 and the reference's 0.5 (Simulation.cpp:477-483).
SHOCK_TUBES = {
    # SodShockTube.hpp:9-31: nx 4, sx 1, 5 levels, t 0.25
    "sod": dict(nelem=128, hi=1.0, t_final=0.25, bc=("specified-state", "specified-state"),
                states=((0.5, 1.0, 0.0, 1.0), (None, 0.125, 0.0, 0.1))),
    # LaxShockTube.hpp: t 0.14
    "lax": dict(nelem=128, hi=1.0, t_final=0.14, bc=("specified-state", "specified-state"),
                states=((0.5, 0.445, 0.698, 3.528), (None, 0.5, 0.0, 0.571))),
    # ShuOsherShock.hpp: nx 16, sx 10, 4 levels, supersonic inflow on the
    # left, outflow on the right, t 1.8; the density ahead of the shock is
    # 1 + 0.2 sin(5 (x - 5))
    "shu_osher": dict(nelem=256, hi=10.0, t_final=1.8, bc=("supersonic-inflow", "supersonic-outflow"),
                      states=((1.0, 3.857, 2.629, 10.333),
                              (None, lambda x: 1.0 + 0.2 * np.sin(5.0 * (x - 5.0)), 0.0, 1.0))),
    # WoodwardColellaBlastWave.hpp: nx 4, 8 levels, slip walls, t 0.038
    "blast": dict(nelem=1024, hi=1.0, t_final=0.038, bc=("slip", "slip"), le=True,
                  states=((0.1, 1.0, 0.0, 1000.0), (0.9, 1.0, 0.0, 0.01), (None, 1.0, 0.0, 100.0))),
    # ModifiedSodShockTube.hpp: nx 4, 5 levels, t 0.2
    "modified_sod": dict(nelem=128, hi=1.0, t_final=0.2, bc=("specified-state", "specified-state"),
                         states=((0.5, 1.0, 0.75, 1.0), (None, 0.125, 0.0, 0.1))),
    # Problem123.hpp (two rarefactions moving apart): nx 4, 5 levels, t 0.15
    "problem123": dict(nelem=128, hi=1.0, t_final=0.15, bc=("specified-state", "specified-state"),
                       states=((0.5, 1.0, -2.0, 0.4), (None, 1.0, 2.0, 0.4))),
    # WoodwardColellaBlastWaveLeft.hpp / ...Right.hpp / ...Collision.hpp: the
    # left and right halves of the blast wave and the collision of the two
    # shocks they send out; nx 4, 5 levels
    "blast_left": dict(nelem=128, hi=1.0, t_final=0.012, bc=("specified-state", "specified-state"),
                       states=((0.5, 1.0, 0.0, 1000.0), (None, 1.0, 0.0, 0.01))),
    "blast_right": dict(nelem=128, hi=1.0, t_final=0.035, bc=("specified-state", "specified-state"),
                        states=((0.5, 1.0, 0.0, 0.01), (None, 1.0, 0.0, 100.0))),
    # (config.json of the collision: alpha_max 0.75)
    "blast_collision": dict(nelem=128, hi=1.0, t_final=0.035, bc=("specified-state", "specified-state"),
                            alpha_max=0.75,
                            states=((0.5, 5.99924, 19.5975, 460.894), (None, 5.99242, -6.19633, 46.0950))),
    # LeBlancShockTube.hpp: nx 16 on [0, 9], 5 levels, gamma 1.66667,
    # alpha_max 1, supersonic outflow at both ends, t 6; the initial
    # condition writes the conservative state itself, (rho, rho u, rho E) =
    # (1, 0, 0.1) for x < 3 and (1e-3, 0, 1e-9) beyond
    "leblanc": dict(nelem=512, hi=9.0, t_final=6.0, gamma=1.66667, alpha_max=1.0, cons=True,
                    bc=("supersonic-outflow", "supersonic-outflow"),
                    states=((3.0, 1.0, 0.0, 0.1), (None, 1e-3, 0.0, 1e-9))),
}


# AcousticWave (TestCases/Euler/1D/AcousticWave: AcousticWave.hpp:9-33,
# order 3, blending, the periodic 4-element segment refined 6 times, t 1.0):
# rho = 1.225 + 0.2 sin(2 pi x) carried at u = M a (M = 0.5, a from p = 1,
# rho = 1.225) through constant pressure; the exact solution is the same
# profile shifted by u t (AcousticWaveExactSolution, lines 35-71).
ACOUSTIC_WAVE = dict(dim=1, order=3, nelem=(256,), lo=(0.0,), hi=(1.0,), gamma=1.4, blending=True,
                     alpha_max=0.5, t_final=1.0)


def acoustic_wave(X, t=0.0, gamma=1.4):
    """AcousticWaveIC at t = 0, its exact solution at t (periodic on [0, 1])."""
    u = 0.5 * np.sqrt(gamma * 1.0 / 1.225)
    x = np.asarray(X[0], float)
    rho = 1.225 + 0.2 * np.sin(2.0 * np.pi * (x - u * t))
    return prim_to_cons(rho, [np.full_like(x, u)], np.ones_like(x), gamma)


def shock_tube(name, order=3):
    """DGSEMOperator keywords of a 1-D case of SHOCK_TUBES, and a function of
    the node coordinates X [1][NE][K] giving its initial state."""
    c = SHOCK_TUBES[name]
    gamma = c.get("gamma", 1.4)

    def prim_at(x):
        x = np.asarray(x, float)
        rho, u, p = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
        done = np.zeros(x.shape, bool)
        for x_end, r, v, q in c["states"]:
            sel = ~done if x_end is None else ~done & ((x <= x_end) if c.get("le") else (x < x_end))
            rho[sel], u[sel], p[sel] = r(x[sel]) if callable(r) else r, v, q
            done |= sel
        return rho, u, p

    def cons_at(x):
        # the states as written: primitive (rho, u, p), or with "cons" the
        # conservative (rho, rho u, rho E) themselves
        a, b, q = prim_at(x)
        if c.get("cons"):
            return np.array([a, b, q])
        return prim_to_cons(a, [b], q, gamma)

    def state(X):
        return cons_at(X[0])

    ends = []
    for x in (0.0, c["hi"]):
        ends.append(tuple(float(v[0]) for v in cons_at(np.array([x])))
                    if c["bc"][len(ends)] in ("specified-state", "supersonic-inflow") else None)
    kw = dict(dim=1, order=order, nelem=(c["nelem"],), lo=(0.0,), hi=(c["hi"],), gamma=gamma,
              blending=True, alpha_max=c.get("alpha_max", 0.5), bc=[c["bc"]], bc_state=[tuple(ends)])
    return kw, state


# The reference's axisymmetric Euler cases (AXISYMMETRIC build, order 1,
# SUBCELL_FV_BLENDING with alpha_max 0.5, fixed dt), on their Gmsh meshes:
# the 15-degree ramp (TestCases/Euler/2D/Ramp: RampIC and RampInletBCVector
# of Ramp.hpp:9-29, Mach 2 free stream rho = 1, a = 1) and Nagashima's ramjet
# (TestCases/Euler/2D/Nagashima_Ramjet/TR0: NagashimaIC and
# NagashimaInletBCVector of Nagashima_Ramjet.hpp:9-42, Mach 2 from p0 =
# 303975 Pa, a0 = 360.63 m/s).  The axis density floor is the one
# Simulation.cpp:925-930 sets from the initial state, 1e-8 rho.
def _nagashima_free_stream(gamma=1.4, R_gas=287.05, Ma=2.0, p0=303975.0, a0=360.63):
    C = 1 + 0.5 * (gamma - 1) * Ma * Ma
    a = a0 / np.sqrt(C)
    ua = Ma * a
    pa = p0 / C ** (gamma / (gamma - 1))
    Ta = (a * a) / (gamma * R_gas)
    rhoa = pa / (R_gas * Ta)
    return np.array([rhoa, rhoa * ua, 0.0, pa / (gamma - 1.0) + 0.5 * rhoa * ua * ua])


# The reference's double Mach reflection (TestCases/Euler/2D/
# DoubleMachReflection: DoubleMachReflection.hpp, config.json): a Mach-10
# shock at 60 degrees to the wall, post-shock state (8, 8 * 7.1447..,
# -8 * 4.125, p = 116.5) ahead of x < 1/6 + y/sqrt(3), quiescent (1.4, 0, 0,
# p = 1) behind; left and bottom1 hold the post-shock state, bottom2 is a
# slip wall, right a supersonic outflow; the top boundary follows the shock:
# the post-shock state where x < 1/6 + (y + 20 t)/sqrt(3).  Order 3,
# blending (alpha_max 0.5, the reference's default), fixed dt 1e-4, one
# refinement of its mesh (par_ref_levels 1), t_final 0.04.
_DMR_U, _DMR_V = 7.144709581221619, 4.125
DMR_POST = np.array([8.0, 8.0 * _DMR_U, -8.0 * _DMR_V,
                     116.5 * 1.0 / (1.4 - 1.0) + 0.5 * 8.0 * (_DMR_U * _DMR_U + _DMR_V * _DMR_V)])
DMR_PRE = np.array([1.4, 0.0, 0.0, 1.0 / (1.4 - 1.0)])
DMR_TOP_SPLIT = dict(normal=(1.0, -1.0 / np.sqrt(3.0)), offset=1.0 / 6.0, speed=20.0 / np.sqrt(3.0),
                     below=DMR_POST, above=DMR_PRE)
# dt 1e-4 is the configured step; from the discontinuous initial state it
# drives the pressure negative at 155 nodes next to the shock foot (1/6, 0) in
# the first stage (alpha_max 0.5 caps the blending), while 5e-5 and 2e-5 keep
# every node positive: dt_start is the step used for starting from the
# initial condition.
DMR = dict(mesh="DoubleMachReflection.msh", order=3, blending=True, alpha_max=0.5, dt=1e-4, dt_start=2e-5,
           t_final=0.04,
           refine=1, bc={"left": ("specified-state", DMR_POST), "bottom1": ("specified-state", DMR_POST),
                         "bottom2": "slip", "right": "supersonic-outflow", "top": ("specified-state", DMR_PRE)},
           split={"top": DMR_TOP_SPLIT})


# BackwardFacingStep (TestCases/Euler/2D/BackwardFacingStep: its mesh, order
# 3, blending, alpha_max 0.5 by default, cfl 1, t 0.01): the post-shock state
# of BackwardFacingStepIC (BackwardFacingStep.hpp:9-37) for x < 0.5, gas at
# rest (rho 1, p 1) beyond; BackwardFacingStepLeftBCVector (line 40) at the
# supersonic inflow, slip walls below and above, supersonic outflow.  Its
# config steps with RK4 ("ode_solver": "RK4"), which a caller composes from
# the operator's Mult.
BFS_INFLOW = np.array([5.9970, 5.9970 * 98.5914, 0.0, 11666.5 / (1.4 - 1.0) + 0.5 * 5.9970 * 98.5914 * 98.5914])
BFS = dict(mesh="BackwardFacingStep.msh", order=3, blending=True, alpha_max=0.5, t_final=0.01,
           bc={"left": ("supersonic-inflow", BFS_INFLOW), "bottom": "slip", "right": "supersonic-outflow",
               "top": "slip"})


def bfs_state(X, gm1i=2.5):
    """BackwardFacingStepIC at the node coordinates X [2][...], with 1/(gamma
    - 1) the config's parameter x1 = 2.5 as written."""
    x = np.asarray(X[0], float)
    left = x < 0.5
    rho = np.where(left, 5.9970, 1.0)
    u = np.where(left, 98.5914, 0.0)
    p = np.where(left, 11666.5, 1.0)
    return np.array([rho, rho * u, np.zeros_like(x), p * gm1i + rho * 0.5 * (u * u)])


def dmr_state(X, gamma=1.4):
    """DoubleMachReflectionIC: the post-shock state where x < 1/6 + y/sqrt(3)."""
    post = X[0] < 1.0 / 6.0 + X[1] / np.sqrt(3.0)
    return np.where(post[None], DMR_POST[:, None, None], DMR_PRE[:, None, None])


RAMP_STATE = np.array([1.0, 2.0, 0.0, (1.0 / 1.4) / (1.4 - 1.0) + 0.5 * 1.0 * 2.0 * 2.0])
RAMP = dict(mesh="Ramp15deg_v2.msh", order=1, blending=True, alpha_max=0.5, dt=1e-4, t_final=3.0,
            state=RAMP_STATE, axis_rho_floor=1e-8 * RAMP_STATE[0],
            bc={"inlet": ("supersonic-inflow", RAMP_STATE), "axis": "axis", "ramp": "slip",
                "outlet": "supersonic-outflow", "top": ("supersonic-inflow", RAMP_STATE)})
NAGASHIMA_STATE = _nagashima_free_stream()
NAGASHIMA = dict(mesh="Nagashima_TR0_10kQuad_v2.msh", order=1, blending=True, alpha_max=0.5, dt=1e-7,
                 t_final=0.02, state=NAGASHIMA_STATE, axis_rho_floor=1e-8 * NAGASHIMA_STATE[0],
                 bc={"inlet": ("supersonic-inflow", NAGASHIMA_STATE), "outlet": "supersonic-outflow",
                     "axis": "axis", "center-body": "slip", "cowl-inner": "slip", "cowl-outer": "slip",
                     "valve": "slip"})


def axisymmetric_state(U, X):
    """The weighted state r U (r = X[1]) of a state U: [4] (uniform) or [4][NE][K]."""
    r = X[1]
    U = np.asarray(U, float)
    if U.ndim == 1:
        U = U[:, None, None] * np.ones_like(r)[None]
    return r[None] * U


def hex_box_mesh(n, lo, hi, warp=0.0):
    """A periodic n[0] x n[1] x n[2] hexahedral mesh of the box [lo, hi]
    (MFEM's vertex order, elements x fastest) whose vertices are moved by
    the smooth periodic warp x_d += warp h_d sin(2 pi x_{d+1} / L_{d+1}), the
    shared vertices of periodic faces identified; returns a HexMesh with
    per-element corners."""
    from .operator import HexMesh
    n = tuple(int(v) for v in n)
    L = [hi[d] - lo[d] for d in range(3)]
    h = [L[d] / n[d] for d in range(3)]
    corners = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))

    def pos(i, j, k):
        p = np.array([lo[0] + i * h[0], lo[1] + j * h[1], lo[2] + k * h[2]])
        q = p.copy()
        for d in range(3):
            e = (d + 1) % 3
            q[d] += warp * h[d] * np.sin(2 * np.pi * (p[e] - lo[e]) / L[e])
        return q

    elem, ec = [], []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                cs = [(i + a, j + b, k + c) for a, b, c in corners]
                elem.append([((c2 % n[2]) * n[1] + (b2 % n[1])) * n[0] + (a2 % n[0]) for a2, b2, c2 in cs])
                ec.append([pos(*c) for c in cs])
    vert = np.zeros((n[0] * n[1] * n[2], 3))
    return HexMesh(vert, np.array(elem), None, None, np.array(ec))


def random_smooth_state(X, dim, seed, amp=0.2, gamma=1.4, rho0=1.0, p0=1.0, L=None):
    """Smooth positive periodic state: a few random Fourier modes on each primitive."""
    rng = np.random.default_rng(seed)
    if L is None:
        L = [X[r].max() - X[r].min() for r in range(dim)]

    def field(scale):
        f = np.zeros_like(X[0])
        for _ in range(3):
            k = rng.integers(1, 3, size=dim)
            ph = rng.uniform(0, 2 * np.pi)
            arg = ph + sum(2 * np.pi * k[r] * X[r] / L[r] for r in range(dim))
            f = f + rng.uniform(-1, 1) * np.sin(arg)
        return scale * f / 3.0

    rho = rho0 * (1.0 + field(amp))
    vel = [field(1.0) for _ in range(dim)]
    p = p0 * (1.0 + field(amp))
    return prim_to_cons(rho, vel, p, gamma)


def kinetic_energy(U, WJ, volume):
    """E_k = (1/|Omega|) sum_e sum_i w_i J_i rho |v|^2 / 2 (LGL quadrature)."""
    dim = U.shape[0] - 2
    ke = 0.5 * sum(U[1 + d] ** 2 for d in range(dim)) / U[0]
    return float(np.sum(WJ * ke) / volume)


def kinetic_energy_rate(U, dU, WJ, volume):
    """dE_k/dt from the RHS: integral of v.d(rho v)/dt - |v|^2/2 drho/dt."""
    dim = U.shape[0] - 2
    rho = U[0]
    vel = [U[1 + d] / rho for d in range(dim)]
    integrand = sum(vel[d] * dU[1 + d] for d in range(dim)) - 0.5 * sum(v * v for v in vel) * dU[0]
    return float(np.sum(WJ * integrand) / volume)


def quadrature_weights(dim, order, nelem, lo, hi):
    """w_i J_i per node, [NE][K]."""
    _, w = lgl_nodes(order)
    Np = order + 1
    h = [(hi[r] - lo[r]) / nelem[r] for r in range(dim)]
    J = h[0] if dim == 1 else h[0] * h[1] if dim == 2 else h[0] * (h[1] * h[2])
    i = np.arange(Np ** dim)
    wt = w[i % Np]
    if dim >= 2:
        wt = wt * w[(i // Np) % Np]
    if dim == 3:
        wt = wt * w[i // (Np * Np)]
    NE = int(np.prod([nelem[r] for r in range(dim)]))
    return np.broadcast_to(wt * J, (NE, Np ** dim)).copy()


# Ensemble of independent 2-D Euler cases: square meshes n x n with n from
# ENSEMBLE_MESHES and degree N from ENSEMBLE_ORDERS; the first half are
# isentropic vortices on [-5,5]^2 with beta ~ U[3,6], the second half
# Kelvin-Helmholtz shear layers on [-1,1]^2 with thickness s ~ U[10,20] and
# blending (alpha_max 0.5).
ENSEMBLE_MESHES = (8, 16, 24, 32, 48, 64)
ENSEMBLE_ORDERS = (2, 3, 4, 5)


def ensemble_cases(ncases=64, seed=20261202, precision="mixed"):
    """Case list drawn from one std::mt19937_64: per case, in this order, a
    mesh index, an order index (index = floor(r * n)) and the case parameter
    (lo + r (hi - lo)).  Returns dicts with the DGSEMOperator keyword
    arguments plus 'kind', 'param' and 'seed' (the KH noise seed)."""
    g = MT19937_64(seed)
    out = []
    for i in range(ncases):
        n = ENSEMBLE_MESHES[int(g.uniform() * len(ENSEMBLE_MESHES))]
        N = ENSEMBLE_ORDERS[int(g.uniform() * len(ENSEMBLE_ORDERS))]
        r = g.uniform()
        c = dict(dim=2, order=N, nelem=(n, n), gamma=1.4, precision=precision)
        if i < ncases // 2:
            c.update(lo=(-5.0, -5.0), hi=(5.0, 5.0), kind="iv", param=3.0 + 3.0 * r, seed=0)
        else:
            c.update(lo=(-1.0, -1.0), hi=(1.0, 1.0), kind="kh", param=10.0 + 10.0 * r,
                     seed=seed + 1 + i, blending=True, alpha_max=0.5)
        out.append(c)
    return out


def ensemble_state(case):
    """Initial state of one ensemble case, [nv][NE][K]."""
    X = node_coords(2, case["order"], case["nelem"], case["lo"], case["hi"])
    if case["kind"] == "iv":
        return isentropic_vortex(X, case["gamma"], beta=case["param"])
    return kelvin_helmholtz(X, case["gamma"], s=case["param"], seed=case["seed"])


def operator_kwargs(case):
    """The DGSEMOperator keyword arguments of an ensemble case."""
    return {k: v for k, v in case.items() if k not in ("kind", "param", "seed")}
