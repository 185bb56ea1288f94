"""Throughput of the DGSEM RHS + SSP-RK3 stage update on one MI355X.

Default workload: 3-D Navier-Stokes Taylor-Green vortex, periodic [0,2pi]^3,
16^3 hexahedra, N = 3 (64^3 nodes x 5 variables), Re = 1600, Ma = 0.1,
Pr = 0.71, BR1 lifting, Chandrashekar volume flux, reference surface flux,
plain FP64.  A "step" is one SSP-RK3 step = three RHS evaluations with the
stage combinations fused in; DOF-updates/s = nodes x RHS evaluations / s.

The state is resident in HBM before the timed region starts.  The JSON line
also carries the roofline of the dominant kernel (k_rhs, timed with HIP
events on its own stream inside the timed region) and a CPU baseline: the
oracle (an FP64 CPU restatement of the reference RHS, not MFEM) timed on this
host on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12      # MI355X dense FP64 (vector = matrix), FLOP/s
PEAK_HBM = 8.0e12        # B/s
METRIC = "RHS DOF-updates/s, plain FP64 vs mixed, 3D TGV N=3; L2 err vs FP64 oracle"

WORKLOADS = {
    "tgv3d_n3": dict(dim=3, order=3, nelem=(16, 16, 16), lo=(0.0,) * 3, hi=(2 * math.pi,) * 3,
                     parabolic=True, mu=1.0 / 1600.0, Pr=0.71, gamma=1.4),
    "tgv3d_n7": dict(dim=3, order=7, nelem=(32, 32, 32), lo=(0.0,) * 3, hi=(2 * math.pi,) * 3,
                     parabolic=True, mu=1.0 / 1600.0, Pr=0.71, gamma=1.4, blending=True, alpha_max=0.1),
    "kh2d_n4": dict(dim=2, order=4, nelem=(64, 64), lo=(-1.0, -1.0), hi=(1.0, 1.0),
                    blending=True, alpha_max=0.5, gamma=1.4),
    "iv2d_n3": dict(dim=2, order=3, nelem=(16, 16), lo=(-5.0, -5.0), hi=(5.0, 5.0), gamma=1.4),
    # the reference's triple-point case: slip walls on all sides, blending
    "triple2d_n3": dict(dim=2, order=3, nelem=(112, 48), lo=(0.0, 0.0), hi=(7.0, 3.0), gamma=1.4,
                        blending=True, alpha_max=0.5, bc=[("slip", "slip"), ("slip", "slip")]),
    # the reference's forward-facing step on its own mesh refined once
    # (13600 quadrilaterals), Mach-3 inflow, outflow, slip walls, blending;
    # the shock is formed by spin-up steps before the timed region
    "ffs2d_n3": dict(dim=2, order=3, mesh="ForwardFacingStep.msh", refine=1, gamma=1.4, blending=True,
                     alpha_max=0.5, spinup=300,
                     bc={"left": ("supersonic-inflow", (1.4, 4.2, 0.0, 1.0 / 0.4 + 0.5 * 1.4 * 9.0)),
                         "right": "supersonic-outflow", "top": "slip", "bottom": "slip"}),
}
# the Taylor-Green workload on a hexahedral mesh of the same box whose
# vertices are moved by a smooth periodic warp (10% of h): the general
# (curvilinear) geometry kernels at the headline size
WORKLOADS["tgv3d_n3_hex"] = dict(WORKLOADS["tgv3d_n3"], hex_warp=0.1)
# the reference's axisymmetric cases on their own meshes (order 1, blending,
# fixed dt; the flow is developed by spin-up steps of that dt)
for _name, _case in (("ramp2d_n1", "RAMP"), ("nagashima2d_n1", "NAGASHIMA")):
    WORKLOADS[_name] = dict(dim=2, order=1, gamma=1.4, axisymmetric=True, spinup=300, case=_case)
# the reference's 1-D Woodward-Colella blast wave (1024 elements of order 3
# between slip walls, blending); spin-up steps at cfl 1 form the waves
WORKLOADS["blast1d_n3"] = dict(dim=1, order=3, nelem=(1024,), lo=(0.0,), hi=(1.0,), gamma=1.4, blending=True,
                               alpha_max=0.5, bc=[("slip", "slip")], bc_state=[(None, None)], spinup=2000)
# the reference's Double Mach reflection on its own mesh refined once
# (par_ref_levels 1: 147456 quadrilaterals), N = 3, blending, with the top
# boundary's moving split state; fixed dt 1e-5 (about cfl 0.5 on this mesh:
# the configured 1e-4 is unstable from the discontinuous start), spin-up
# steps of that dt before the timed region
WORKLOADS["dmr2d_n3"] = dict(dim=2, order=3, gamma=1.4, spinup=300, case="DMR", dt=1e-5)
MESH_DIR = os.path.join(ROOT, "tests", "golden", "meshes")
# 64 independent 2-D Euler cases (isentropic vortices and blended
# Kelvin-Helmholtz layers, meshes 8^2..64^2, N = 2..5) advanced together
ENSEMBLE = "ensemble64"


def resolve(cfg):
    """A workload's settings, those of a mesh case from prandtl_amd.cases
    (the axisymmetric cases, DMR); the workload's own dt, if it has one,
    takes precedence."""
    if "case" not in cfg:
        return cfg
    from prandtl_amd import cases
    c = getattr(cases, cfg["case"])
    out = dict(cfg, mesh=c["mesh"], order=c["order"], blending=c["blending"], alpha_max=c["alpha_max"],
               bc=c["bc"], dt=cfg.get("dt", c["dt"]))
    for k in ("state", "axis_rho_floor", "split", "refine"):
        if k in c:
            out[k] = c[k]
    del out["case"]
    return out


def quad_mesh(cfg):
    from prandtl_amd.operator import QuadMesh
    return QuadMesh.read(os.path.join(MESH_DIR, cfg["mesh"]), refine=cfg.get("refine", 0))


def axi_kwargs(cfg):
    return dict(axisymmetric=True, axis_rho_floor=cfg["axis_rho_floor"]) if cfg.get("axisymmetric") else {}


def hex_mesh(cfg):
    from prandtl_amd import cases
    return cases.hex_box_mesh(cfg["nelem"], cfg["lo"], cfg["hi"], warp=cfg["hex_warp"])


def box_kwargs(cfg):
    return {k: v for k, v in cfg.items() if k not in ("hex_warp", "dim", "nelem", "lo", "hi")}


def op_kwargs(cfg):
    """A box workload's operator settings (without its spin-up count)."""
    return {k: v for k, v in cfg.items() if k != "spinup"}


def make_operator(cfg, quirks, precision):
    from prandtl_amd import DGSEMOperator
    if "hex_warp" in cfg:
        return DGSEMOperator.hex(hex_mesh(cfg), cfg["order"], quirks=quirks, precision=precision,
                                 **{k: v for k, v in box_kwargs(cfg).items() if k != "order"})
    if "mesh" in cfg:
        op = DGSEMOperator.quad(quad_mesh(cfg), cfg["order"], bc=cfg["bc"], gamma=cfg["gamma"],
                                blending=cfg.get("blending", False), alpha_max=cfg.get("alpha_max", 0.5),
                                quirks=quirks, precision=precision, **axi_kwargs(cfg))
        for name, sp in cfg.get("split", {}).items():
            op.set_boundary_split(name, sp["normal"], sp["offset"], sp["speed"], sp["below"], sp["above"])
        return op
    return DGSEMOperator(quirks=quirks, precision=precision, **op_kwargs(cfg))


def make_oracle(cfg, quirks, threads):
    """The oracle on the same workload (test infrastructure, CPU baseline and op counts)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    if "hex_warp" in cfg:
        m = hex_mesh(cfg)
        d = dict(vert=m.vert, elem=m.elem, elem_coords=m.elem_coords)
        return O.Oracle.hex(d, cfg["order"], quirks=quirks, threads=threads,
                            **{k: v for k, v in box_kwargs(cfg).items() if k != "order"})
    if "mesh" in cfg:
        m = quad_mesh(cfg)
        nattr = int(m.bdr_attr.max())
        kinds, states = ["slip"] * nattr, [[1.0, 0.0, 0.0, 2.5]] * nattr
        for key, v in cfg["bc"].items():
            a = m.names[key]
            kinds[a - 1] = v if isinstance(v, str) else v[0]
            if not isinstance(v, str):
                states[a - 1] = list(v[1])
        d = dict(vert=m.vert, elem=m.elem, elem_coords=m.elem_coords, bdr=m.bdr, bdr_attr=m.bdr_attr,
                 attr_bc=kinds, attr_state=states)
        orc = O.Oracle.quad(d, cfg["order"], gamma=cfg["gamma"], blending=cfg.get("blending", False),
                            alpha_max=cfg.get("alpha_max", 0.5), quirks=quirks, threads=threads,
                            **axi_kwargs(cfg))
        for name, sp in cfg.get("split", {}).items():
            orc.set_boundary_split(m.names[name], sp["normal"], sp["offset"], sp["speed"], sp["below"],
                                   sp["above"])
        return orc
    return O.Oracle(threads=threads, quirks=quirks, **op_kwargs(cfg))


def initial_state(name, cfg):
    import numpy as np
    from prandtl_amd import cases
    if "mesh" in cfg:
        # the reference's ForwardFacingStepIC, RampIC, NagashimaIC: the free
        # stream everywhere (r U in axisymmetric mode)
        m = quad_mesh(cfg)
        K = (cfg["order"] + 1) ** 2
        if "split" in cfg:
            # DoubleMachReflectionIC: the post-shock state left of the shock line
            return cases.dmr_state(m.node_coords(cfg["order"])).reshape(4, -1)
        U = np.asarray(cfg["state"] if "state" in cfg else cfg["bc"]["left"][1], float)
        if cfg.get("axisymmetric"):
            return cases.axisymmetric_state(U, m.node_coords(cfg["order"])).reshape(4, -1)
        return np.repeat(U[:, None], m.nelem * K, axis=1)
    if "hex_warp" in cfg:
        X = hex_mesh(cfg).node_coords(cfg["order"])
    else:
        X = cases.node_coords(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    if name.startswith("tgv"):
        return cases.taylor_green(X)
    if name.startswith("kh"):
        return cases.kelvin_helmholtz(X)
    if name.startswith("triple"):
        return cases.triple_point(X)
    if name.startswith("blast"):
        return cases.shock_tube("blast")[1](X)
    return cases.isentropic_vortex(X)


def algorithmic_counts(name, quirks):
    """Algorithmic FP64 op counts of one RHS (oracle op-counting instantiation
    on the initial state, committed as a fixture): per node for a single
    case, totals over all cases for the ensemble (quirks None)."""
    path = os.path.join(ROOT, "tests", "golden", "flop_counts.json")
    with open(path) as f:
        table = json.load(f)
    if quirks is None:
        return table[name]
    return table["%s_%s" % (name, "faithful" if quirks else "corrected")]


def host_info():
    """CPU model and core counts of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def host_threads():
    """Threads for the all-cores baseline: the job's CPU share where the
    environment states it (OMP_NUM_THREADS), else every core of the host."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return os.cpu_count() or 1


def median_rate(orc, u, nodes, reps, per_rep):
    """Median over reps repetitions of per_rep RHS evaluations (after a warm-up)."""
    orc.Mult(u)
    rates, spent = [], 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(per_rep):
            orc.Mult(u)
        el = time.perf_counter() - t0
        spent += el
        rates.append(nodes * per_rep / el)
    rates.sort()
    return rates[len(rates) // 2], spent


def cpu_baseline(name, cfg, quirks, u, budget_s):
    """Oracle RHS on this host's cores, bounded sample of the same workload:
    the median of 5 repetitions on all the job's cores and on one core.  The
    repetition length is sized from one timed evaluation so that both legs
    together take about budget_s."""
    info = host_info()
    nthr = host_threads()
    out = {}
    for label, threads in (("all", nthr), ("one", 1)):
        orc = make_oracle(cfg, quirks, threads)
        nodes = orc.NE * orc.K
        orc.Mult(u)   # the first evaluation pays for first-touch allocation
        t0 = time.perf_counter()
        orc.Mult(u)
        one = max(time.perf_counter() - t0, 1e-6)
        share = budget_s * (0.4 if label == "all" else 0.6)
        per_rep = max(1, min(200, int(share / 5 / one)))
        rate, spent = median_rate(orc, u, nodes, 5, per_rep)
        out[label] = (rate, per_rep, spent, threads)
    rate, per_rep, spent, threads = out["all"]
    r1, p1, s1, _ = out["one"]
    return {"value": rate, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": "median of 5 repetitions of %d RHS evaluations of %s (%d nodes) on %d threads "
                      "(OpenMP over elements/faces; the job's CPU share, OMP_NUM_THREADS) in %.1f s; "
                      "CPU restatement of the reference RHS (not MFEM)" % (per_rep, name, nodes, threads, spent),
            "one_thread": {"value": r1, "sample": "median of 5 repetitions of %d RHS evaluations in %.1f s"
                                                  % (p1, s1)},
            "host": info}


def cpu_baseline_ensemble(case_list, budget_s):
    """Oracle RHS over the ensemble's cases in order, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from prandtl_amd import cases
    threads = host_threads()
    nodes, n, el = 0, 0, 0.0
    for c in case_list:
        kw = cases.operator_kwargs(c)
        kw.pop("precision")
        orc = O.Oracle(threads=threads, **kw)
        u = cases.ensemble_state(c)
        t0 = time.perf_counter()
        orc.Mult(u)
        el += time.perf_counter() - t0
        nodes += orc.NE * orc.K
        n += 1
        if el > budget_s:
            break
    return {"value": nodes / el, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": "one RHS evaluation of each of the first %d ensemble cases (%d nodes) in %.1f s, "
                      "OpenMP over elements/faces" % (n, nodes, el), "host": host_info()}


def run_ensemble(args):
    """Aggregate throughput of the ensemble launch vs sequential per-case launches."""
    import torch
    from prandtl_amd import cases
    from prandtl_amd.operator import Ensemble

    precision = args.precision or "mixed"
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    case_list = cases.ensemble_cases(64, 20261202, precision=precision)
    ens = Ensemble(case_list)
    us, dts = [], []
    for c, op in zip(case_list, ens.ops):
        u = torch.tensor(cases.ensemble_state(c).reshape(-1), dtype=torch.float64, device=dev)
        z = torch.empty_like(u)
        op.Mult(u, z)
        # fixed dt from CFL 0.5 of the initial wave speed
        dts.append(0.5 * op.hmin / (op.GetMaxCharSpeed(reset=True) * (c["order"] + 1) ** 2))
        us.append(u)
    nodes = sum(op.NE * op.K for op in ens.ops)
    # launch form: one workgroup per tile, or persistent with g workgroups
    # per CU; the fastest first stage on these states is kept
    form, form_us = (0, None) if args.no_autotune else ens.autotune(us, dts)
    peaks = None if args.no_peaks else measured_peaks(dev)

    def ens_steps(k):
        for _ in range(k):
            ens.Step(us, dts)

    def seq_steps(k):
        for _ in range(k):
            for op, u, dt in zip(ens.ops, us, dts):
                op.Step(u, dt)

    ens_steps(args.warmup)
    seq_steps(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens_steps(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    t0 = time.perf_counter()
    seq_steps(args.steps)
    torch.cuda.synchronize()
    elapsed_seq = time.perf_counter() - t0

    # k_ens duration: each replay of the step graph is three k_ens launches
    # back to back on this stream, bracketed by events
    psteps = min(args.steps, 50)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(psteps)]
    for a, b in evs:
        a.record(stream)
        ens.Step(us, dts)
        b.record(stream)
    torch.cuda.synchronize()
    t_launch = sum(a.elapsed_time(b) for a, b in evs) / psteps / 3 * 1e-3
    ok = all(op.check_state(u) for op, u in zip(ens.ops, us))

    counts = algorithmic_counts(ENSEMBLE, None)
    nv = 4
    flops_launch = (counts["volume"] + counts["surface"] + counts["fv"] + counts["indicator"]
                    + counts["assembly"] + 4 * nv * nodes)
    bytes_launch = 8 * nv * (1 + 1 + 2.0 / 3.0) * nodes
    kname = "k_ens" if form == 0 else "k_ens_q"
    roof = with_measured(roofline(flops_launch, bytes_launch, t_launch, kname + " (8 variants N=2..5 x blending)",
                                  measured_traffic(ENSEMBLE, kname)), peaks)
    value = nodes * 3 * args.steps / elapsed
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f64 (mixed form; the Euler path has no FP32 stage)",
        "data": "synthetic (analytic initial conditions, seeded case draws)",
        "config": {"workload": ENSEMBLE + ": 64 2-D Euler cases, 32 isentropic vortices + 32 blended "
                               "Kelvin-Helmholtz, meshes {8..64}^2, N 2..5, seed 20261202",
                   "cases": len(case_list), "nodes": nodes, "vars": nv, "rhs_per_step": 3,
                   "precision": precision, "state_ok": bool(ok), "hip_graph": True,
                   "launch_form": {"workgroups_per_cu": form or None,
                                   "stage_us_by_form": [round(t, 2) for t in form_us] if form_us else None},
                   "sequential_per_case": {"value": nodes * 3 * args.steps / elapsed_seq,
                                           "ms_per_step": elapsed_seq / args.steps * 1e3,
                                           "launches_per_step": 3 * len(case_list)},
                   "speedup_vs_sequential": elapsed_seq / elapsed, "measured_peaks": peaks},
        "roofline": roof,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_ensemble(case_list, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    ens.close()


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel, from the committed FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/hbm_traffic_<workload>.txt,
    calibrated against a copy kernel); None when there is none."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "traffic" and parts[1] == workload and parts[2] == kernel:
                return float(parts[3])
    return None


def mixed_comparison(name, cfg, u_host, dt, steps, warmup, autotune, dev):
    """The mixed form on the same workload, in the same process: its
    throughput, and its RHS and kinetic-energy rate against the plain FP64
    GPU RHS at the initial state (the FP64 RHS matches the oracle to 1e-12
    relative max-norm, tests/test_gpu_parity.py)."""
    import numpy as np
    import torch
    from prandtl_amd import DGSEMOperator, cases
    opd = DGSEMOperator(precision="fp64", quirks=False, **cfg)
    opm = DGSEMOperator(precision="mixed", quirks=False, **cfg)
    u0 = torch.tensor(u_host.reshape(-1), dtype=torch.float64, device=dev)
    zd, zm = torch.empty_like(u0), torch.empty_like(u0)
    opd.Mult(u0, zd)
    opm.Mult(u0, zm)
    torch.cuda.synchronize()
    rd, rm = zd.cpu().numpy(), zm.cpu().numpy()
    WJ = cases.quadrature_weights(cfg["dim"], cfg["order"], cfg["nelem"], cfg["lo"], cfg["hi"])
    vol = float(np.prod([cfg["hi"][r] - cfg["lo"][r] for r in range(cfg["dim"])]))
    ek_d = cases.kinetic_energy_rate(u_host, rd.reshape(u_host.shape), WJ, vol)
    ek_m = cases.kinetic_energy_rate(u_host, rm.reshape(u_host.shape), WJ, vol)
    shapes = opm.autotune(u0)[0] if autotune else None
    u = u0.clone()
    for _ in range(warmup):
        opm.Step(u, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opm.Step(u, dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nodes = opm.NE * opm.K
    return {"value": nodes * 3 * steps / el, "ms_per_step": el / steps * 1e3, "launch_shapes": shapes,
            "state_ok": bool(opm.check_state(u)),
            "rhs_rel_l2_vs_fp64": float(np.linalg.norm(rm - rd) / np.linalg.norm(rd)),
            "dEk_dt_fp64": ek_d, "dEk_dt_mixed": ek_m,
            "dEk_dt_rel_diff": abs(ek_m - ek_d) / abs(ek_d)}


def measured_peaks(dev):
    """Achievable peaks of this GPU, measured before the timed region
    (prandtl_peak_fma, prandtl_peak_copy): FP64 and unpacked FP32 FMA chains,
    and device copies of 1 GiB (HBM) and of 27 MB (54 MB of traffic, the
    RHS kernel's, which the 256 MB Infinity Cache can hold)."""
    import ctypes
    import torch
    from prandtl_amd._lib import lib
    L = lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    sink = torch.zeros(1, dtype=torch.float64, device=dev)
    out = {}
    for bits in (64, 32):
        tf = ctypes.c_double()
        rc = L.prandtl_peak_fma(bits, ctypes.c_void_p(sink.data_ptr()), ctypes.byref(tf), stream)
        assert rc == 0, "prandtl_peak_fma failed (%d)" % rc
        out["fp%d_tflops" % bits] = tf.value
    for label, nbytes in (("1GiB", 1 << 30), ("27MB", 27000000 // 16 * 16)):
        a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        us = ctypes.c_double()
        rc = L.prandtl_peak_copy(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), nbytes, 20,
                                 ctypes.byref(us), stream)
        assert rc == 0, "prandtl_peak_copy failed (%d)" % rc
        out["copy_GBps_%s" % label] = 2 * nbytes / us.value / 1e3
        del a, b
    torch.cuda.synchronize()
    return out


def with_measured(roof, peaks):
    """The roofline entry's achieved rate as a fraction of the measured peak
    of its unit as well (the contract's peak is the datasheet one)."""
    if peaks:
        m = peaks["copy_GBps_1GiB"] if roof["unit"] == "GB/s" else peaks["fp64_tflops"]
        roof["measured_peak"] = m
        roof["frac_of_measured"] = roof["achieved"] / m
    return roof


def roofline(flops_launch, bytes_launch, t_launch, kernel, traffic=None):
    """Roofline entry of one kernel from its algorithmic counts per launch."""
    ai = flops_launch / bytes_launch
    if ai >= PEAK_FP64 / PEAK_HBM:
        roof = {"bound": "mfma", "achieved": flops_launch / t_launch / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "pipe": "fp64 (vector and matrix peaks are equal on MI355X)"}
    else:
        roof = {"bound": "hbm", "achieved": bytes_launch / t_launch / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = traffic
    roof["kernel"] = kernel
    roof["avg_launch_us"] = t_launch * 1e6
    roof["algorithmic_flops_per_launch"] = flops_launch
    roof["algorithmic_bytes_per_launch"] = bytes_launch
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="tgv3d_n3", choices=sorted(WORKLOADS) + [ENSEMBLE])
    ap.add_argument("--faithful", action="store_true",
                    help="reproduce the reference's defects (its NS-3D path is unstable on this case)")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default: fp64 for single cases, mixed for the ensemble")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes of both kernels")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the FMA-chain and copy measurements of the machine peaks")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-form comparison of a plain FP64 Navier-Stokes run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the path is single-GPU in this build "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.workload == ENSEMBLE:
        run_ensemble(args)
        return

    import torch
    from prandtl_amd import DGSEMOperator

    args.precision = args.precision or "fp64"
    name = args.workload
    cfg = resolve(dict(WORKLOADS[name]))
    quirks = bool(args.faithful)
    dev = torch.device("cuda:0")
    # a non-default stream (default priority): the RK3 step is captured once
    # as a hipGraph on it and replayed
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    op = make_operator(cfg, quirks, args.precision)
    u_host = initial_state(name, cfg)
    u = torch.tensor(u_host.reshape(-1), dtype=torch.float64, device=dev)
    nodes = op.NE * op.K
    if cfg.get("spinup"):
        # untimed: develop the flow (shock and blending) from the uniform start
        if "dt" in cfg:
            op.Run(u, 0.0, 1e30, dt=cfg["dt"], max_steps=cfg["spinup"])
        else:
            op.Run(u, 0.0, 1e30, cfl=1.0, max_steps=cfg["spinup"])
        torch.cuda.synchronize()

    # fixed dt from CFL 0.1 of the initial wave speed (TGV config policy)
    z = torch.empty_like(u)
    op.Mult(u, z)
    lam = op.GetMaxCharSpeed()
    # (the axisymmetric cases: their configured fixed dt)
    dt = cfg["dt"] if "dt" in cfg else 0.1 * op.hmin / (lam * (cfg["order"] + 1) ** 2)

    # launch shapes of both kernels: timed on this state, fastest kept
    shape, shape_us = ({"rhs": 0, "visc": 0}, None) if args.no_autotune else op.autotune(u)
    peaks = None if args.no_peaks else measured_peaks(dev)

    for _ in range(args.warmup):
        op.Step(u, dt)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.Step(u, dt)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # per-kernel durations: the same steps again, every launch bracketed by
    # HIP events on the stream the kernels run on
    psteps = min(args.steps, 50)
    op.profile_begin()
    tp = time.perf_counter()
    for _ in range(psteps):
        op.Step(u, dt)
    torch.cuda.synchronize()
    elapsed_prof = time.perf_counter() - tp
    prof = op.profile_end()
    ok = op.check_state(u)

    rhs_evals = 3 * args.steps
    value = nodes * rhs_evals / elapsed
    ms_rhs, n_rhs = prof["rhs"]
    ms_visc, n_visc = prof["visc"]
    t_rhs = ms_rhs / max(n_rhs, 1) * 1e-3

    counts = algorithmic_counts(name, quirks)
    nv = op.nv
    # Algorithmic FP64 flops per node (oracle op counts of the reference
    # formulas).  The reference computes subcell FV fluxes for every element
    # and multiplies them by alpha = 0; those count only for the elements that
    # blend (alpha > 0) at the timed state.
    alpha, _ = op.alpha()
    blend_frac = float((alpha > 0).mean()) if cfg.get("blending") else 0.0
    fl_rhs = (counts["volume"] + counts["surface"] + counts["fv"] * blend_frac + counts["indicator"]
              + counts["assembly"] + 4 * nv)          # k_rhs, with the RK stage combination
    fl_visc = counts["lifting"] + counts["viscous"]    # k_visc
    # Compulsory HBM bytes per node of one RHS + RK stage: read the stage
    # state, read u^n on two stages of three, write the stage output.  The
    # viscous intermediates and geometry tables are implementation bytes
    # (measured, in "traffic").
    by_node = 8 * nv * (1 + 1 + 2.0 / 3.0)
    kname = "k_rhs<%d,%d>" % (cfg["dim"], cfg["order"] + 1)
    traffic = measured_traffic(name, kname) if (args.precision == "fp64" and not quirks) else None
    roof = with_measured(roofline(fl_rhs * nodes, by_node * nodes, t_rhs, kname, traffic), peaks)
    t_visc = ms_visc / max(n_visc, 1) * 1e-3
    kernels = {kname: {"avg_launch_us": t_rhs * 1e6, "flops_per_node": fl_rhs,
                       "tflops": fl_rhs * nodes / t_rhs / 1e12,
                       "frac_fp64": fl_rhs * nodes / t_rhs / PEAK_FP64}}
    if n_visc:
        vname = "k_visc<%d,%d>" % (cfg["dim"], cfg["order"] + 1)
        kernels[vname] = {"avg_launch_us": t_visc * 1e6, "flops_per_node": fl_visc,
                          "tflops": fl_visc * nodes / t_visc / 1e12,
                          "frac_fp64": fl_visc * nodes / t_visc / PEAK_FP64,
                          "traffic": measured_traffic(name, vname) if (args.precision == "fp64" and not quirks)
                          else None}
    # the whole RHS at the measured DOF rate
    fl_node = fl_rhs + (fl_visc if n_visc else 0.0)
    whole = {"flops_per_dof": fl_node, "bytes_per_dof": by_node,
             "tflops": value * fl_node / 1e12, "frac_fp64": value * fl_node / PEAK_FP64,
             "hbm_GBps": value * by_node / 1e9, "frac_hbm": value * by_node / PEAK_HBM,
             "bound": "fp64" if fl_node / by_node >= PEAK_FP64 / PEAK_HBM else "hbm",
             "blend_frac": blend_frac}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (FP32 BR1/viscous)",
        "data": "synthetic (analytic initial condition)",
        "config": {"workload": name + (" (reference defects reproduced)" if quirks else
                                      " (reference defects corrected)"),
                   "elements": op.NE, "order": cfg["order"], "nodes": nodes, "vars": nv,
                   "rhs_per_step": 3, "dt": dt, "state_ok": bool(ok), "hip_graph": True,
                   "launch_shapes": shape, "launch_shape_us": shape_us,
                   "ms_per_step_event_profiled": elapsed_prof / psteps * 1e3,
                   "kernel_ms": {"k_rhs": ms_rhs / max(n_rhs, 1), "k_visc": ms_visc / max(n_visc, 1)},
                   "kernels": kernels, "rhs_roofline": whole, "measured_peaks": peaks},
        "roofline": roof,
    }
    if args.precision == "fp64" and cfg.get("parabolic") and not quirks and not args.no_mixed \
            and "hex_warp" not in cfg:
        m = mixed_comparison(name, cfg, u_host, dt, args.steps, args.warmup, not args.no_autotune, dev)
        m["speedup_vs_fp64"] = m["value"] / value
        line["config"]["mixed"] = m
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(name, cfg, quirks, u_host, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
