"""Pins of the oracle's time-dependent outer states (a straight discontinuity
moving at constant speed, the form of the reference's
DoubleMachReflectionTopBCFunction, DoubleMachReflection.hpp:33-57, evaluated
per boundary point at the operator's time as SpecifiedStateBdrFaceIntegrator.cpp:
24-34 does) and of the stage times of an RK3 step (RK3SSPSolver's SetTime at
t, t + dt, t + dt/2).

- Far past / far future: every boundary point is on one side, so the RHS
  equals the RHS with that side's constant state, bitwise.
- In between: boundary nodes clearly on one side carry exactly that side's
  constant-state RHS; interior nodes away from the boundary are unchanged.
- The split moves with time: the node set that takes the "below" state
  grows as speed t grows.
- An RK3 step equals three RHS evaluations at t, t + dt, t + dt/2 combined
  by the Shu-Osher stages (a step with every stage at t differs).
"""
import numpy as np

import oracle_lib as O
import quad_meshes as Q
from prandtl_amd import cases

G = 1.4
BELOW = np.array([2.0, 1.0, -0.5, 8.0])
ABOVE = np.array([1.0, 0.2, 0.0, 2.6])


def mesh_with(state_top):
    m = Q.grid(6, 3, hi=(3.0, 1.0), periodic=(False, False))
    # attributes: 1 left, 2 right, 3 bottom, 4 top
    m["attr_bc"] = ["supersonic-outflow", "supersonic-outflow", "slip", "specified-state"]
    m["attr_state"] = np.array([ABOVE, ABOVE, ABOVE, state_top])
    return m


def oracle(state_top, split=None, order=3):
    orc = O.Oracle.quad(mesh_with(state_top), order)
    if split is not None:
        orc.set_boundary_split(4, *split)
    return orc


def state(orc):
    X = orc.coords()
    rho = 1.0 + 0.2 * np.sin(2 * X[0]) * np.cos(3 * X[1])
    u = 0.3 + 0.1 * np.cos(X[0])
    v = 0.1 * np.sin(X[1])
    p = 1.0 + 0.1 * np.cos(X[0] + X[1])
    return cases.prim_to_cons(rho, [u, v], p, G)


SPLIT = ((1.0, -0.5), 1.0, 5.0, BELOW, ABOVE)   # below where x - y/2 - 1 - 5 t < 0


def test_split_far_past_and_future_equal_constant_states():
    orc = oracle(ABOVE, SPLIT)
    u = state(orc)
    for t, const in ((-10.0, ABOVE), (10.0, BELOW)):
        orc.set_time(t)
        r = orc.Mult(u)
        ref = oracle(const).Mult(u)
        assert np.array_equal(r, ref), t


def test_split_in_between_and_moving():
    orc = oracle(ABOVE, SPLIT)
    u = state(orc)
    X = orc.coords()
    ra = oracle(ABOVE).Mult(u).reshape(u.shape)
    rb = oracle(BELOW).Mult(u).reshape(u.shape)
    top = np.abs(X[1] - 1.0) < 1e-12
    prev = None
    for t in (0.0, 0.1, 0.2):
        orc.set_time(t)
        r = orc.Mult(u).reshape(u.shape)
        side = X[0] - 0.5 * X[1] - 1.0 - 5.0 * t
        # rows of nodes touching the top boundary: which state they see
        below = top & (side < -1e-9)
        above = top & (side > 1e-9)
        assert below.any() and above.any()
        assert np.array_equal(r[:, below], rb[:, below])
        assert np.array_equal(r[:, above], ra[:, above])
        # nodes of elements away from the top keep the same RHS
        inner = X[1] < 0.6
        assert np.array_equal(r[:, inner], ra[:, inner])
        if prev is not None:
            assert below.sum() > prev   # the discontinuity moves to larger x
        prev = below.sum()


def test_rk3_stage_times():
    dt = 0.02
    orc = oracle(ABOVE, SPLIT)
    u = state(orc).reshape(-1)
    # at t0 the split crosses the top boundary at x = 1.85, at t0 + dt at 1.95:
    # the node at x = 1.862 changes sides within the step
    t0 = 0.07
    orc.set_time(t0)
    step = orc.Step(u.copy(), dt)
    assert abs(orc.time - (t0 + dt)) <= 1e-15
    # the same step by hand: stages at t, t + dt, t + dt/2
    man = oracle(ABOVE, SPLIT)

    def L(y, t):
        man.set_time(t)
        return man.Mult(y).reshape(-1)

    y = u + dt * L(u, t0)
    y = y + dt * L(y, t0 + dt)
    y = 3.0 / 4 * u + 1.0 / 4 * y
    y = y + dt * L(y, t0 + dt / 2)
    x = 1.0 / 3 * u + 2.0 / 3 * y
    assert np.array_equal(step, x)
    # every stage at t0 instead: a different step (the split moves 5 dt)
    y = u + dt * L(u, t0)
    y = 3.0 / 4 * u + 1.0 / 4 * (y + dt * L(y, t0))
    x0 = 1.0 / 3 * u + 2.0 / 3 * (y + dt * L(y, t0))
    assert not np.array_equal(step, x0)
