"""Full-length runs of the benchmark cases, with their physical diagnostics.

 - run_tgv: 3-D Navier-Stokes Taylor-Green vortex, 16^3 hexahedra, N = 3,
   Re = 1600, Ma = 0.1, Pr = 0.71, to t = 2 with a fixed dt from CFL 0.1 of
   the initial wave speed (the dt policy of the reference's TGV setup,
   TestCases/NavierStokes/3D/TaylorGreenVortex/config.json:27), the last step
   clipped to t_final (Simulation.cpp:1007-1022).  Samples the kinetic energy
   E_k = <rho |v|^2 / 2> and its rate dE_k/dt from the RHS.
 - run_kh: 2-D Kelvin-Helmholtz layer, 64^2 quads, N = 4, subcell blending
   (alpha_max 0.5), to t = 3 at CFL 0.5 with the reference's variable-dt
   policy.  Samples positivity and the blending coefficients.

Every step runs on the GPU through DGSEMOperator; numpy only reduces the
sampled states on the host.
"""
import json
import math
import sys

import numpy as np

from . import cases
from .operator import DGSEMOperator, TimeLoop

TGV = dict(dim=3, order=3, nelem=(16, 16, 16), lo=(0.0,) * 3, hi=(2 * math.pi,) * 3,
           parabolic=True, mu=1.0 / 1600.0, Pr=0.71, gamma=1.4, quirks=False)
KH = dict(dim=2, order=4, nelem=(64, 64), lo=(-1.0, -1.0), hi=(1.0, 1.0), gamma=1.4,
          blending=True, alpha_max=0.5)


def _device_state(u, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(u).reshape(-1), dtype=torch.float64, device=dev)


def run_tgv(precision="fp64", t_final=2.0, cfl=0.1, samples=20, dev="cuda:0"):
    """Returns {"t", "ek", "dekdt", "steps", "dt", "state_ok"}: samples of E_k
    and dE_k/dt at evenly spaced steps, first at t = 0, last at t_final."""
    import torch
    cfg = TGV
    op = DGSEMOperator(precision=precision, **cfg)
    X = cases.node_coords(3, cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    u0 = cases.taylor_green(X)
    WJ = cases.quadrature_weights(3, cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    vol = float(np.prod([cfg["hi"][r] - cfg["lo"][r] for r in range(3)]))
    stream = torch.cuda.Stream(device=dev)
    out = {"t": [], "ek": [], "dekdt": []}
    with torch.cuda.stream(stream):
        u = _device_state(u0, dev)
        z = torch.empty_like(u)

        def sample(t):
            op.Mult(u, z)
            torch.cuda.synchronize()
            U = u.cpu().numpy().reshape(u0.shape)
            dU = z.cpu().numpy().reshape(u0.shape)
            out["t"].append(t)
            out["ek"].append(cases.kinetic_energy(U, WJ, vol))
            out["dekdt"].append(cases.kinetic_energy_rate(U, dU, WJ, vol))

        op.Mult(u, z)
        dt = cfl * op.hmin / (op.GetMaxCharSpeed() * (cfg["order"] + 1) ** 2)
        nsteps = int(math.ceil(t_final / dt - 1e-8))
        every = max(1, nsteps // samples)
        sample(0.0)

        def cb(t, steps, done):
            if done or steps % every == 0:
                sample(t)

        t, steps = TimeLoop(op, t_final, dt=dt).run(u, callback=cb)
        ok = op.check_state(u)
    torch.cuda.synchronize()
    out.update(steps=steps, dt=dt, t_final=t, state_ok=bool(ok), precision=precision)
    return out


def run_kh(precision="fp64", t_final=3.0, cfl=0.5, samples=12, dev="cuda:0"):
    """Returns {"t", "min_rho", "min_p", "alpha_max", "blended", "steps",
    "state_ok"}: per sample the smallest density and pressure, the largest
    alpha and the fraction of elements with alpha > 0."""
    import torch
    cfg = KH
    op = DGSEMOperator(precision=precision, **cfg)
    X = cases.node_coords(2, cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    u0 = cases.kelvin_helmholtz(X, cfg["gamma"])
    out = {"t": [], "min_rho": [], "min_p": [], "alpha_max": [], "blended": []}
    state_ok = True
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        u = _device_state(u0, dev)

        def sample(t):
            nonlocal state_ok
            state_ok = state_ok and op.check_state(u)
            U = u.cpu().numpy().reshape(u0.shape)
            rho = U[0]
            p = (cfg["gamma"] - 1.0) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / rho)
            a, _ = op.alpha()
            out["t"].append(t)
            out["min_rho"].append(float(rho.min()))
            out["min_p"].append(float(p.min()))
            out["alpha_max"].append(float(a.max()))
            out["blended"].append(float((a > 0).mean()))

        sample(0.0)
        marks = [t_final * (k + 1) / samples for k in range(samples)]

        def cb(t, steps, done):
            while marks and (t >= marks[0] - 1e-12 or done):
                marks.pop(0)
                sample(t)
                if done:
                    marks.clear()

        t, steps = TimeLoop(op, t_final, cfl=cfl, variable_dt=True).run(u, callback=cb)
    torch.cuda.synchronize()
    out.update(steps=steps, t_final=t, state_ok=bool(state_ok), precision=precision)
    return out


def main(argv):
    which = argv[1] if len(argv) > 1 else "tgv"
    prec = argv[2] if len(argv) > 2 else "fp64"
    r = run_tgv(prec) if which == "tgv" else run_kh(prec)
    print(json.dumps(r))


if __name__ == "__main__":
    main(sys.argv)
