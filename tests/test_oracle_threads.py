"""The oracle's RHS, blending coefficients and BR1 gradients do not depend on
its thread count, bit for bit.  Every loop the threads share either writes
disjoint nodes or runs on one thread: element loops, the interior faces of a
box one direction at a time (a node takes at most one face term per
direction, so its terms are added in the serial order), and the faces of
general meshes and all boundary faces on one thread.  The CPU baseline of the
benchmark runs the oracle on all host cores, so its results must be those of
the single-threaded parity runs."""
import numpy as np
import pytest

import oracle_lib as O
import quad_meshes as Q
from prandtl_amd import cases


def smooth_state(X, dim):
    x = X[0]
    y = X[1] if dim > 1 else 0.0 * x
    rho = 1.0 + 0.2 * np.sin(2 * x + 0.3) * np.cos(y + 0.1)
    vel = [0.3 * np.cos(x + 2 * y), -0.2 * np.sin(3 * x - y), 0.1 * np.cos(x - y)][:dim]
    p = 1.0 + 0.1 * np.cos(x) * np.sin(2 * y + 0.2)
    return cases.prim_to_cons(rho, vel, p, 1.4).reshape(-1)


def box3(threads):
    return O.Oracle(3, 3, (3, 4, 2), (0, 0, 0), (1.0, 1.2, 0.8), parabolic=True, mu=2e-2, Pr=0.71,
                    quirks=True, blending=True, alpha_max=0.5, threads=threads)


def box2_walls(threads):
    bc = [("periodic", "periodic"), ("no-slip-adiabatic", "no-slip-adiabatic")]
    st = [(None, None), ((0.0,) * 5, (0.1, 0.0, 0.0, 0.0, 0.0))]
    return O.Oracle(2, 4, (3, 4), (0, 0), (1.0, 1.0), parabolic=True, mu=0.02, Pr=0.71, quirks=False,
                    bc=bc, bc_state=st, threads=threads)


def quad(threads):
    return O.Oracle.quad(Q.grid(4, 3, perturb=0.2), 3, blending=True, parabolic=True, mu=0.01, Pr=0.71,
                         threads=threads)


@pytest.mark.parametrize("make,dim", [(box3, 3), (box2_walls, 2), (quad, 2)], ids=["box3", "box2_walls", "quad"])
def test_oracle_is_thread_count_independent(make, dim):
    one, many = make(1), make(7)
    u = smooth_state(one.coords(), dim)
    r1, r7 = one.Mult(u), many.Mult(u)
    assert np.array_equal(r1, r7)
    assert np.array_equal(one.alpha(), many.alpha())
    assert np.array_equal(one.gradients(), many.gradients())
    assert one.GetMaxCharSpeed() == many.GetMaxCharSpeed()
