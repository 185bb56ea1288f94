"""Full-length runs of the benchmark cases on the GPU (prandtl_amd/physics.py).

 - TGV 16^3 N=3 to t = 2: the mixed form's dissipation rate -dE_k/dt stays
   within 1e-4 relative of the plain FP64 run's at every sample (the FP64
   RHS equals the oracle's to 1e-12, tests/test_gpu_parity.py), E_k within
   1e-6; E_k decreases and the dissipation rate stays positive.
 - KH 64^2 N=4 with blending to t = 3 at CFL 0.5: density and pressure stay
   positive, alpha stays within [0, alpha_max] and some elements blend.
"""
import json

import pytest

from prandtl_amd import physics

pytestmark = pytest.mark.gpu


def test_tgv_to_t2_mixed_matches_fp64(cuda):
    d = physics.run_tgv("fp64")
    m = physics.run_tgv("mixed")
    print(json.dumps({"fp64": d, "mixed": m}), flush=True)
    assert d["state_ok"] and m["state_ok"]
    assert d["steps"] == m["steps"] and abs(d["t_final"] - 2.0) < 1e-12
    worst = max(abs(a - b) / abs(b) for a, b in zip(m["dekdt"], d["dekdt"]))
    worst_ek = max(abs(a - b) / abs(b) for a, b in zip(m["ek"], d["ek"]))
    print("TGV t=2: steps %d, max rel. diff of -dEk/dt mixed vs fp64 %.3e, of Ek %.3e"
          % (d["steps"], worst, worst_ek), flush=True)
    assert worst <= 1e-4
    assert worst_ek <= 1e-6
    assert all(b < a for a, b in zip(d["ek"], d["ek"][1:]))
    assert all(r < 0 for r in d["dekdt"])


def test_kh_to_t3_stays_physical(cuda):
    r = physics.run_kh("fp64")
    print(json.dumps(r), flush=True)
    assert r["state_ok"] and abs(r["t_final"] - 3.0) < 1e-12
    assert min(r["min_rho"]) > 0 and min(r["min_p"]) > 0
    assert max(r["alpha_max"]) <= 0.5
    assert max(r["blended"]) > 0
