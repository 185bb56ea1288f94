"""Pins of the oracle's faithful-mode branches (the reference's defects as
written) and of the blending coefficient at intermediate eta.

Each pin computes a faithful-minus-corrected difference of the oracle's own
RHS, chosen so that everything but one defect cancels, and compares it with
that defect restated from the reference's lines through the separately
pinned point functions (volume_flux, face_flux, viscous_flux, D^, D^2).
Each pin is then run once more with the matching misreading planted in the
oracle (orc_params.plant) and must fail there:

 - A-3, DGSEMIntegrator.cpp:620: the viscous volume flux of every node takes
   the stale state of node K-1 (the member id1 left by the previous element
   loop or by ComputeFVFluxes' zeta sweep), with the node's own (faithful,
   A-8-aliased) gradients.
   Isolated as [RHS(mu) - RHS(0)]_faithful - [RHS(mu) - RHS(0)]_corrected
   (constant mu: the viscous terms are linear in mu; the face viscous fluxes
   use the true states in both modes).  Plant: node 0's state.
 - A-2, DGSEMIntegrator.cpp:702: H_inviscid(id2).SetCol(k, g) puts the last
   eta-pair flux g of node (i, j, m) into column m of node (i, j, k) for every
   zeta pair m < k.  Isolated at mu = 0 on states with rho = 1 and w = 0,
   where A-1, A-3 and A-8 vanish.  Plants: h itself (no defect), and g
   taken from the xi loop.
 - A-8, DGSEMIntegrator.cpp:467-481 with BasicOperations.cpp:33-34: the 2-D
   Navier-Stokes face flux reads (rho, v, E), the velocities ComputePressure
   left in the momentum slots.  Isolated at mu = 0 in 2-D.  Plant: no alias.
 - A-8 in EntropyGrad2PrimGrad (BasicOperations.cpp:137-138, 164): the
   density row multiplies velocities, so grad rho gains
   (rho/p)(1 - rho) v.grad v and grad p gains p/rho times that, in closed
   form.  Plant: momenta.
 - alpha, DGSEMOperator.cpp:25, 60-80: on element-local states whose rho p
   has known shifted-Legendre modes, eta and the logistic alpha take their
   closed-form values at intermediate eta.  Plant: threshold from N + 1.

A-9 (the lifting flux of Entropy2Conserv's aliased w) is pinned in closed
form in tests/test_oracle_ns_walls.py.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from prandtl_amd import cases

G = 1.4
PR = 0.71


def box(dim, order, nelem, hi):
    h = [hi[r] / nelem[r] for r in range(dim)]
    J = h[0] if dim == 1 else h[0] * h[1] if dim == 2 else h[0] * (h[1] * h[2])
    nrm = [h[1], h[0]] if dim == 2 else [h[1] * h[2], h[0] * h[2], h[0] * h[1]]
    return h, J, np.array(nrm)


def line_sum(M, F, dim, Np):
    """sum_l M[i_r, l] F[.., l, .., r] for every node, summed over r:
    F has shape [NE, (Np,)*dim (k, j, i), nv, dim]."""
    if dim == 2:
        return (np.einsum("il,ejlv->ejiv", M, F[..., 0]) + np.einsum("jl,eliv->ejiv", M, F[..., 1]))
    return (np.einsum("il,ekjlv->ekjiv", M, F[..., 0]) + np.einsum("jl,ekliv->ekjiv", M, F[..., 1])
            + np.einsum("kl,eljiv->ekjiv", M, F[..., 2]))


def to_vars(A, nv, NE, K):
    """[NE, (Np,)*dim, nv] -> [nv, NE, K]"""
    return np.moveaxis(A.reshape(NE, K, nv), 2, 0)


# ---------------------------------------------------------------- A-3

def a3_residual(plant):
    dim, order, nelem, hi, mu = 3, 3, (2, 2, 2), (1.0, 1.2, 0.9), 0.05
    Np, K, nv = order + 1, (order + 1) ** 3, 5
    kw = dict(parabolic=True, Pr=PR, threads=4)
    orcs = {(q, m): O.Oracle(dim, order, nelem, (0.0,) * dim, hi, mu=m, quirks=q,
                             plant=plant if q else 0, **kw)
            for q in (True, False) for m in (mu, 0.0)}
    X = orcs[(False, mu)].coords()
    u = cases.random_smooth_state(X, dim, 41, L=list(hi))
    R = {k: o.Mult(u).reshape(nv, -1, K) for k, o in orcs.items()}
    got = (R[(True, mu)] - R[(True, 0.0)]) - (R[(False, mu)] - R[(False, 0.0)])
    # primitive gradients [dim][nv][NE][K]: the faithful ones carry A-8's
    # aliased grad rho and grad p, which the stale state's p/rho no longer
    # cancels in the heat flux
    Gf, Gc = orcs[(True, mu)].gradients(), orcs[(False, mu)].gradients()
    NE = Gc.shape[2]
    _, J, nrm = box(dim, order, nelem, hi)
    _, Dh, _ = O.dmat(order)
    U = u.reshape(nv, NE, K)
    dF = np.zeros((NE, K, nv, dim))
    for e in range(NE):
        stale = U[:, e, K - 1]
        for n in range(K):
            dqf = [Gf[c, :, e, n] for c in range(dim)]
            dqc = [Gc[c, :, e, n] for c in range(dim)]
            dF[e, n] = (O.viscous_flux(dim, G, PR, mu, stale, dqf) - O.viscous_flux(dim, G, PR, mu, U[:, e, n], dqc)) * nrm
    want = to_vars(line_sum(Dh, dF.reshape(NE, Np, Np, Np, nv, dim), dim, Np) / J, nv, NE, K)
    scale = np.abs(want).max()
    assert scale > 1e-3 * np.abs(R[(False, mu)] - R[(False, 0.0)]).max()
    return float(np.abs(got - want).max() / scale)


def test_a3_stale_state_in_viscous_volume_flux():
    r = a3_residual(0)
    print("A-3 residual", r)
    assert r <= 1e-11


def test_a3_pin_fails_on_node0_misreading():
    r = a3_residual(O.PLANT_STALE_NODE0)
    print("A-3 residual with node 0 planted", r)
    assert r > 1e-3


# ---------------------------------------------------------------- A-2

def a2_residual(plant):
    dim, order, nelem, hi = 3, 3, (2, 2, 2), (1.0, 1.2, 0.9)
    Np, K, nv = order + 1, (order + 1) ** 3, 5
    kw = dict(parabolic=True, Pr=PR, mu=0.0, threads=4)
    of = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=True, plant=plant, **kw)
    oc = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=False, **kw)
    X = of.coords()
    r = cases.random_smooth_state(X, dim, 43, L=list(hi))
    vel = [r[1] / r[0], r[2] / r[0], np.zeros_like(r[0])]
    p = (G - 1.0) * (r[4] - 0.5 * r[0] * (vel[0] ** 2 + vel[1] ** 2 + (r[3] / r[0]) ** 2))
    u = cases.prim_to_cons(np.ones_like(p), vel, p, G)   # rho = 1, w = 0
    got = of.Mult(u).reshape(nv, -1, K) - oc.Mult(u).reshape(nv, -1, K)
    NE = got.shape[1]
    _, J, nrm = box(dim, order, nelem, hi)
    _, _, Dh2 = O.dmat(order)
    U = u.reshape(nv, NE, K)
    my, mz = np.array([0.0, nrm[1], 0.0]), np.array([0.0, 0.0, nrm[2]])
    node = lambda i, j, k: (k * Np + j) * Np + i
    want = np.zeros_like(got)
    for e in range(NE):
        for k in range(Np):
            for j in range(Np):
                for i in range(Np):
                    acc = np.zeros(nv)
                    for m in range(k):
                        nm = node(i, j, m)
                        # g: node (i, j, m)'s last eta flux (with (i, N, m)); 0 when j = N
                        g = (O.volume_flux(dim, G, U[:, e, nm], U[:, e, node(i, Np - 1, m)], my, my)[0]
                             if j < Np - 1 else np.zeros(nv))
                        h = O.volume_flux(dim, G, U[:, e, nm], U[:, e, node(i, j, k)], mz, mz)[0]
                        acc += Dh2[k, m] * (g - h)
                    want[:, e, node(i, j, k)] = -acc / J
    scale = np.abs(want).max()
    assert scale > 1e-3 * np.abs(oc.Mult(u)).max()
    return float(np.abs(got - want).max() / scale)


def test_a2_zeta_partner_takes_last_eta_flux():
    r = a2_residual(0)
    print("A-2 residual", r)
    assert r <= 1e-12


@pytest.mark.parametrize("plant", [O.PLANT_ZETA_F, O.PLANT_ZETA_XI])
def test_a2_pin_fails_on_misreadings(plant):
    r = a2_residual(plant)
    print("A-2 residual with plant", plant, r)
    assert r > 1e-3


# ---------------------------------------------------------------- A-8 (face)

def a8_face_residual(plant):
    dim, order, nelem, hi = 2, 3, (3, 2), (1.0, 0.8)
    Np, K, nv = order + 1, (order + 1) ** 2, 4
    kw = dict(parabolic=True, Pr=PR, mu=0.0, threads=4)
    of = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=True, plant=plant, **kw)
    oc = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=False, **kw)
    X = of.coords()
    u = cases.random_smooth_state(X, dim, 47, L=list(hi))
    got = of.Mult(u).reshape(nv, -1, K) - oc.Mult(u).reshape(nv, -1, K)
    NE = got.shape[1]
    _, J, nrm = box(dim, order, nelem, hi)
    _, w = O.lgl(order)
    U = u.reshape(nv, NE, K)
    alias = lambda s: np.concatenate([s[:1], s[1:1 + dim] * (1.0 / s[0]), s[1 + dim:]])
    want = np.zeros_like(got)
    nx = nelem[0]
    for e in range(NE):
        ex, ey = e % nx, e // nx
        for r in range(dim):
            # this element's high face: it is the face's first element (normal +e_r)
            en = ((ex + 1) % nx + ey * nx) if r == 0 else (ex + ((ey + 1) % nelem[1]) * nx)
            nor = np.zeros(dim)
            nor[r] = nrm[r]
            for l in range(Np):
                n1 = (l * Np + (Np - 1)) if r == 0 else ((Np - 1) * Np + l)
                n2 = (l * Np) if r == 0 else l
                s1, s2 = U[:, e, n1], U[:, en, n2]
                dF = (O.face_flux(dim, G, 1, alias(s1), alias(s2), nor)[0] - O.face_flux(dim, G, 1, s1, s2, nor)[0])
                want[:, e, n1] -= dF / (w[0] * J)
                want[:, en, n2] += dF / (w[0] * J)
    scale = np.abs(want).max()
    assert scale > 1e-3 * np.abs(oc.Mult(u)).max()
    return float(np.abs(got - want).max() / scale)


def test_a8_ns_face_flux_reads_aliased_velocities():
    r = a8_face_residual(0)
    print("A-8 face residual", r)
    assert r <= 1e-12


def test_a8_face_pin_fails_without_alias():
    r = a8_face_residual(O.PLANT_FACE_NO_ALIAS)
    print("A-8 face residual with the alias removed", r)
    assert r > 1e-3


# ---------------------------------------------------------------- A-8 (gradients)

def a8_grad_residual(plant):
    dim, order, nelem, hi = 3, 3, (2, 2, 2), (1.0, 1.2, 0.9)
    kw = dict(parabolic=True, Pr=PR, mu=0.01, threads=4)
    of = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=True, plant=plant, **kw)
    oc = O.Oracle(dim, order, nelem, (0.0,) * dim, hi, quirks=False, **kw)
    X = of.coords()
    u = cases.random_smooth_state(X, dim, 53, L=list(hi))
    u[0] *= 1.5                                   # rho away from 1: the alias matters
    of.Mult(u)
    oc.Mult(u)
    gf, gc = of.gradients(), oc.gradients()
    U = u.reshape(5, *gc.shape[2:])
    rho = U[0]
    vel = U[1:4] / rho
    p = (G - 1.0) * (U[4] - 0.5 * rho * (vel ** 2).sum(axis=0))
    worst = 0.0
    for c in range(dim):
        assert np.abs(gf[c, 1:4] - gc[c, 1:4]).max() <= 1e-12 * np.abs(gc[c, 1:4]).max()
        drho = (rho / p) * (1.0 - rho) * (vel * gc[c, 1:4]).sum(axis=0)
        s = np.abs(drho).max()
        worst = max(worst, np.abs((gf[c, 0] - gc[c, 0]) - drho).max() / s,
                    np.abs((gf[c, 4] - gc[c, 4]) - p / rho * drho).max() / np.abs(p / rho * drho).max())
    return worst


def test_a8_gradient_alias_closed_form():
    r = a8_grad_residual(0)
    print("A-8 gradient residual", r)
    assert r <= 1e-9


def test_a8_gradient_pin_fails_without_alias():
    r = a8_grad_residual(O.PLANT_GRAD_NO_ALIAS)
    print("A-8 gradient residual with the alias removed", r)
    assert r > 1e-3


# ---------------------------------------------------------------- alpha

def alpha_residual(plant, N=4):
    nel = (3, 3)
    orc = O.Oracle(2, N, nel, (0.0, 0.0), (1.0, 1.0), blending=True, alpha_max=1.0, plant=plant)
    Np = N + 1
    x1, _ = O.lgl(N)
    i = np.arange(Np * Np)
    xl, yl = x1[i % Np], x1[i // Np]                     # element-local coordinates
    P = lambda n, t: np.polynomial.legendre.legval(2.0 * t - 1.0, [0.0] * n + [1.0])
    T = 0.5 * 10.0 ** (-1.8 * N ** 0.25)                 # DGSEMOperator.cpp:25
    NE = nel[0] * nel[1]
    # eta ~ f T per element (the mean mode carries energy 2^2 = 4)
    f = np.linspace(0.7, 1.35, NE)
    a = np.sqrt(4.0 * T * f)                             # degree (N, 0) mode
    b = np.where(np.arange(NE) % 3 == 0, np.sqrt(4.0 * 0.9 * T), 0.0)   # degree (N-1, N-1) mode
    p = 2.0 + a[:, None] * P(N, xl)[None, :] + b[:, None] * (P(N - 1, xl) * P(N - 1, yl))[None, :]
    u = cases.prim_to_cons(np.ones_like(p), [np.zeros_like(p)] * 2, p, G)
    orc.Mult(u)
    mm, m1, m2 = 4.0 + a ** 2 + b ** 2, 4.0 + b ** 2, 4.0
    eta = np.maximum(1.0 - m1 / mm, 1.0 - m2 / m1)
    al = 1.0 / (1.0 + np.exp(-9.21024 * (eta - T) / T))
    al = np.where(al < 0.001, 0.0, np.where(al > 0.999, 1.0, al))
    assert 0.05 < al.min() and al.max() < 0.98           # intermediate: no clip
    e_err = np.abs(orc.eta() - eta).max() / eta.max()
    return max(e_err, float(np.abs(orc.alpha() - al).max()))


def test_alpha_at_intermediate_eta():
    r = alpha_residual(0)
    print("alpha residual", r)
    assert r <= 1e-9


def test_alpha_pin_fails_on_threshold_from_n_plus_1():
    r = alpha_residual(O.PLANT_THRESH_NP1)
    print("alpha residual with N+1 planted", r)
    assert r > 1e-2
