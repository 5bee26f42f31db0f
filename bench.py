"""Benchmark: velocity solves/s at 1024² (DIVA, MG-PCG), mixed precision vs FP64, on one MI355X.

A step is one full velocity solve (update_velocities!: driving stress, Picard with DIVA coefficient updates and
multigrid-preconditioned CG, velocity-profile reconstruction) of the 1024 × 1024 floating-channel workload
(the DIVA shelf, 512 km square at dx = 500 m), from u = v = 0 and zero DIVA state, with every input already in HBM.
`value` is the mixed-precision form's solves/s; the FP64 form's solves/s and the mixed result's relative L2 error
against it are reported beside it.

Roofline: the three HBM-bound operator/smoother kernels of the headline solve — the finest-grid l1-Jacobi sweep
(k_sweep_m), the PCG operator (k_apply_mt<double>) and the Galerkin stencil sweep on the first coarse level
(k_gp_sweep_t)
— each timed with HIP events on the solver's own stream (nk_solver_options.profile_kernel) during one extra solve;
`roofline` is the one with the most device time per solve, the others are listed under roofline.kernels with their
algorithmic bytes per launch. (The LDS-resident coarse tail, k_g_tail, is latency-bound on one workgroup and has no
HBM roofline.)

CPU baseline: the CPU oracle (plain C, one thread) timed on a bounded sample of the same workload — one DIVA
coefficient update and three MG-PCG iterations at 1024² — and scaled to one solve with the FP64 GPU run's Picard
and CG counts.

Serving (bounded sample): --serve requests of the seeded serving stream (sides 64–1024) replayed at their Poisson arrival times
through nk_server (--serve-slots workers); p50/p99 of finish − arrival and solves/s over the run, beside serial FP64
GPU solves/s of the first --serve-serial requests.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nunatak.jl_amd"))

from nunatak_hip import abi, workloads  # noqa: E402
from nunatak_hip.model import HIPModel  # noqa: E402

METRIC = "velocity solves/s @1024² mixed vs FP64 at rel err≤1e-8; served p50/p99 ms"
HBM_PEAK_GBS = 8000.0
STATE = ("U", "V", "ETASIGMA", "F2", "BETABU", "BETABV", "UB", "VB")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make(N, precision, profile=0, check_interval=8, picard=None, graphs=1, coarse=abi.COARSE_REDISCRETIZED):
    g, phys, opts, fields = workloads.diva_shelf(N=N, dx=512e3 / N, precision=precision)
    opts.coarse_operator = coarse
    opts.profile_kernel = profile
    opts.graphs = graphs
    opts.check_interval = check_interval
    if picard:
        opts.picard = abi.fixed_iterations(picard)
    m = HIPModel(g, phys, opts)
    workloads.load_into(m, fields)
    return m, (g, phys, opts, fields)


def reset_state(m):
    for name in STATE:
        m.fill(name, 0.0)


def timed_solves(m, warmup, steps, label):
    infos = []
    for k in range(warmup):
        reset_state(m)
        info = m.update_velocities(geometry=True)
        log(f"[bench] {label} warmup {k + 1}/{warmup}: {info['elapsed_ms']:.1f} ms, picard {info['picard_iterations']}, "
            f"cg {info['total_linear_iterations']}")
    m.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        reset_state(m)
        infos.append(m.update_velocities(geometry=True))
    m.synchronize()
    wall = time.perf_counter() - t0
    for k, info in enumerate(infos):
        log(f"[bench] {label} step {k + 1}/{steps}: {info['elapsed_ms']:.1f} ms, picard {info['picard_iterations']}, "
            f"cg {info['total_linear_iterations']}, refinement {info['refinement_steps']}")
    return wall, infos


def cpu_baseline(setup, fp64_info):
    """Oracle (C, 1 thread) on a bounded sample of the workload, scaled to one solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    g, phys, opts, fields = setup
    o_opts = abi.default_options(preconditioner=abi.PRECONDITIONER_MULTIGRID, linear=abi.fixed_iterations(3))
    o = oracle.OracleModel(g, phys, o_opts)
    workloads.load_into(o, fields)
    o.update_geometry()
    o.compute_driving_stress()
    t0 = time.perf_counter()
    o.update_coefficients()
    t_coef = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.solve()
    t_solve3 = time.perf_counter() - t0
    o_opts0 = abi.default_options(preconditioner=abi.PRECONDITIONER_MULTIGRID, linear=abi.fixed_iterations(0))
    o0 = oracle.OracleModel(g, phys, o_opts0)
    workloads.load_into(o0, fields)
    o0.update_geometry()
    o0.compute_driving_stress()
    o0.update_coefficients()
    t0 = time.perf_counter()
    o0.solve()
    t_setup = time.perf_counter() - t0
    t_iter = max(t_solve3 - t_setup, 0.0) / 3.0
    picard = fp64_info["picard_iterations"]
    cg = fp64_info["total_linear_iterations"]
    t_solve = picard * (t_coef + t_setup) + cg * t_iter
    log(f"[bench] cpu oracle: coefficient update {t_coef:.2f} s, solve setup {t_setup:.2f} s, CG iteration {t_iter:.2f} s "
        f"→ {t_solve:.1f} s per solve ({picard} Picard, {cg} CG iterations)")
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle at 1024²: one DIVA coefficient update ({t_coef:.2f} s), MG-PCG setup ({t_setup:.2f} s) and "
                      f"3 MG-PCG iterations ({t_iter:.2f} s each), scaled by the FP64 GPU solve's {picard} Picard and "
                      f"{cg} CG iterations"}


def serving_leg(n, rate_hz, slots, serial_sample):
    """Config 5 on a bounded sample: n requests of the stated stream (Poisson arrivals at rate_hz, sides log-uniform
    in [64, 1024] rounded to 16) through the server; the serial FP64 baseline solves the first serial_sample
    requests one after another on fresh FP64 handles."""
    from nunatak_hip.server import serve
    reqs = workloads.serving_stream(n=n, rate_hz=rate_hz, seed=5)
    log(f"[bench] serving {n} requests at {rate_hz} /s on {slots} slots")
    out = serve(reqs, slots=slots, progress=max(1, n // 10))
    t_serial = 0.0
    for k, (_, g, p, o, f) in enumerate(reqs[:serial_sample]):
        o64 = abi.default_options(linear=o.linear, picard=o.picard, preconditioner=o.preconditioner,
                                  precision=abi.PRECISION_FP64, coarse_operator=abi.COARSE_REDISCRETIZED)
        m = HIPModel(g, p, o64)
        workloads.load_into(m, f)
        t0 = time.perf_counter()
        m.update_velocities(geometry=True)
        t_serial += time.perf_counter() - t0
        m.close()
    serial = serial_sample / t_serial if serial_sample else None
    sides = [r[1].Nx for r in reqs]
    log(f"[bench] served: p50 {out['p50_ms']:.0f} ms, p99 {out['p99_ms']:.0f} ms, {out['solves_per_s']:.2f} solves/s; "
        f"serial FP64 {serial:.2f} solves/s on the first {serial_sample}")
    return {"requests": n, "rate_hz": rate_hz, "slots": slots, "sides_min_max": [min(sides), max(sides)],
            "p50_ms": out["p50_ms"], "p99_ms": out["p99_ms"], "solves_per_s": out["solves_per_s"],
            "serial_fp64_solves_per_s": serial, "serial_sample": serial_sample,
            "workload": "serving stream (seed 5): SSA grounded/floating channel per request, mixed precision "
                        "(FP32 Galerkin V-cycle), MG-PCG linear 1e-8, Picard 1e-6"}


def transport_leg(N, steps):
    """Thickness transport at N² on the device: `steps` split Runge-Kutta steps (cache, then three stages of tendency
    and substep) with prescribed velocities, upwind face thickness and sources, no geometry update between stages, so
    the time is the transport kernels' (each entry point synchronises). Algorithmic bytes per step: three tendencies
    (u, v, H, a, m, φ, mᵍ read, G written: 8 planes) and three substeps (H⁻, G, Hᶜ read, H, Hᶜ written: 5 planes) and
    the cache copy (2 planes), 8 B per cell each."""
    L = N * 1e3
    g = abi.default_grid(N, N, 1e3, 1e3, topology=(abi.BOUNDED, abi.PERIODIC), halo=4, Nsigma=2)
    m = HIPModel(g, abi.default_physics(masking=0, grounding=abi.GROUNDING_SUBGRID), abi.default_options())
    x = abi.node_coordinates(g, "H")[0]
    xu = abi.node_coordinates(g, "U")[0]
    m.set("H", 100 + 500 * np.exp(-((x - L / 4) / (L / 16)) ** 2))
    m.set("BED", np.full_like(x, -300.0))
    m.set("U", 1e-5 * np.sin(np.pi * xu / L))
    m.set("SMB", np.full_like(x, 3e-9))
    m.set("BASAL_MELT", 1e-9 * x / L)
    m.update_geometry()
    dt = 0.5 * m.cell_advection_timescale()

    def step():
        m.cache_thickness()
        for k, beta in enumerate((3.0, 2.0, 1.0)):
            m.compute_thickness_tendency(abi.ADVECTION_UPWIND1)
            m.thickness_substep(dt / beta, k == 2)

    step()
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    m.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    bytes_step = 8.0 * N * N * (3 * 8 + 3 * 5 + 2)
    log(f"[bench] transport {N}x{N}: {ms:.3f} ms per RK3 step, {bytes_step / (ms * 1e-3) / 1e9:.0f} GB/s (algorithmic, wall)")
    m.close()
    return {"workload": f"thickness transport {N}x{N}, (Bounded, Periodic), UpwindBiased(order = 1), SplitRungeKutta3, "
                        "prescribed velocities, sources, no geometry update", "ms_per_step": ms,
            "algorithmic_bytes_per_step": bytes_step, "GB_s_wall": bytes_step / (ms * 1e-3) / 1e9, "steps": steps}


def energy_leg(N, steps, Ns=10):
    """Enthalpy step at N² × Nσ on the device: strain heating, then nk_energy_step (layer fluxes, basal state, column
    system with explicit upwind transport, tridiagonal solve, drainage, basal water, thermal diagnostics), on a sheet
    with flow, temperate layers and melt. Algorithmic bytes per step (8 B per value): strain heating reads u, v and
    ησ and writes Φ (Nσ + 2 planes + Nσ); the assembly reads E, uσ, vσ, Φ (4Nσ planes) and 11 planes (H, Hⁿ, φ, m,
    mᵍ, Hw, G, Tₛ, Tₒ, two drag pairs counted as one each) and writes a, b, c, f, W and the basal state (5Nσ + 5); the
    solve reads a, b, c, f and the basal state, H, φ, Hw (4Nσ + 7) and writes E, t, T, ω, A, Hw, mᵍ, Hⁿ (5Nσ + 3)."""
    rng = np.random.default_rng(9)
    g = abi.default_grid(N, N, 1e3, 1e3, topology=(abi.BOUNDED, abi.PERIODIC), halo=4, Nsigma=Ns)
    phys = abi.default_physics(masking=0, grounding=abi.GROUNDING_SUBGRID, rate_factor_depth=abi.RATE_FACTOR_3D)
    m = HIPModel(g, phys, abi.default_options())
    x = abi.node_coordinates(g, "H")[0]
    H = 500 + 1500 * np.exp(-((x - N * 500.0) / (N * 300.0)) ** 2)
    m.set("H", H)
    m.set("HN", H * 0.999)
    m.set("BED", np.full_like(x, -200.0))
    Epm = 2009.0 * (273.15 - 223.15)
    E = Epm - 2009.0 * 30 * rng.random((Ns, N, N))
    E[rng.random(E.shape) < 0.2] = Epm + 0.02 * 3.34e5
    m.set("ENTHALPY", E)
    for name, val in (("GEOTHERMAL", 0.06), ("TSURF", 250.0), ("TOCEAN", 271.15), ("BASAL_MELT", 1e-9)):
        m.set(name, np.full_like(x, val))
    m.set("ETASIGMA", np.full((Ns, N, N), 1e13))
    xu = abi.node_coordinates(g, "U")[0]
    m.set("U", 1e-5 * np.sin(np.pi * xu / (N * 1e3)))
    m.set("USIGMA", np.broadcast_to(1e-5 * np.sin(np.pi * xu / (N * 1e3)), (Ns,) + xu.shape).copy())
    m.set("BETAU", np.full_like(xu, 1e8))
    m.set("UB", 5e-6 * np.ones_like(xu))
    m.update_geometry()

    def step():
        m.compute_strain_heating()
        m.energy_step(3e7, False)

    step()
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    m.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    planes = (2 * Ns + 2) + (4 * Ns + 11) + (5 * Ns + 5) + (4 * Ns + 7) + (5 * Ns + 3)
    bytes_step = 8.0 * N * N * planes
    log(f"[bench] energy {N}x{N}x{Ns}: {ms:.3f} ms per step, {bytes_step / (ms * 1e-3) / 1e9:.0f} GB/s (algorithmic, wall)")
    m.close()
    return {"workload": f"enthalpy step {N}x{N}, Nsigma={Ns}, (Bounded, Periodic), flow, temperate layers, melt",
            "ms_per_step": ms, "algorithmic_bytes_per_step": bytes_step, "GB_s_wall": bytes_step / (ms * 1e-3) / 1e9,
            "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--picard", type=int, default=None, help="FixedIterations(n) Picard (profiling runs only)")
    ap.add_argument("--graphs", type=int, default=1, help="0: launch kernel by kernel (for external profilers)")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--stream-curve", action="store_true",
                    help="device-to-device copy bandwidth from 16 MB to 4 GB (the 256 MB Infinity Cache shows in it); "
                         "prints one JSON line and exits")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: one device copy of 1 GiB (torch) as the counter calibration, then one "
                         "Picard iteration of the mixed solve launched kernel by kernel; prints nothing else")
    ap.add_argument("--coarse", choices=["galerkin", "rediscretized"], default="galerkin",
                    help="coarse operators of the mixed form's FP32 V-cycle (the FP64 form keeps the reference's)")
    ap.add_argument("--serve", type=int, default=48, help="requests in the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-serial", type=int, default=12, help="requests timed serially in FP64 for the baseline")
    ap.add_argument("--transport", type=int, default=20, help="RK3 thickness steps timed in the transport leg (0: skip)")
    ap.add_argument("--transport-only", action="store_true",
                    help="run only the thickness-transport and enthalpy legs; prints their JSON objects")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the velocity solve of one grid is not "
              "sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    N, K, Wm = args.size, args.steps, args.warmup
    if args.transport_only:
        print(json.dumps({"transport": transport_leg(N, max(args.transport, 1)), "energy": energy_leg(N, 10)}), flush=True)
        return
    os.environ.setdefault("NK_VERBOSE", "1")  # one library progress line per Picard iteration
    if args.stream_curve:
        import torch
        curve = []
        for mb in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096):
            n = mb * (1 << 20) // 4
            a = torch.ones(n, dtype=torch.float32, device="cuda")
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = max(5, min(200, 4096 // mb))
            e0.record()
            for _ in range(reps):
                b.copy_(a)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            gbs = 2 * mb * (1 << 20) / (ms * 1e-3) / 1e9
            curve.append({"MB": mb, "ms": ms, "GB_s": gbs})
            log(f"[bench] copy {mb:5d} MB: {ms:8.3f} ms, {gbs:7.0f} GB/s (read + write)")
            del a, b
        print(json.dumps({"stream_copy_curve": curve}), flush=True)
        return
    if args.traffic_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] traffic probe: copied 1 GiB")
        m, _ = make(N, abi.PRECISION_MIXED, picard=1, graphs=0, coarse=abi.COARSE_GALERKIN)
        reset_state(m)
        info = m.update_velocities(geometry=True)
        log(f"[bench] traffic probe: 1 Picard iteration, {info['total_linear_iterations']} CG iterations")
        return

    log(f"[bench] building FP64 and mixed handles at {N}x{N}")
    results = {}
    models = {}
    for label, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
        if args.only and args.only != label:
            continue
        coarse = abi.COARSE_GALERKIN if (label == "mixed" and args.coarse == "galerkin") else abi.COARSE_REDISCRETIZED
        m, setup = make(N, prec, picard=args.picard, graphs=args.graphs, coarse=coarse)
        models[label] = (m, setup)
        wall, infos = timed_solves(m, Wm, K, label)
        results[label] = {"wall": wall, "infos": infos, "u": m.get("U"), "v": m.get("V")}

    out = {"metric": METRIC, "unit": "solves/s", "n_gpus": 1, "steps": K, "warmup": Wm, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "data": "synthetic (seeded numpy PCG64 generator of the DIVA shelf)"}
    if "fp64" in results and "mixed" in results:
        a = np.concatenate([results["mixed"]["u"].ravel(), results["mixed"]["v"].ravel()])
        b = np.concatenate([results["fp64"]["u"].ravel(), results["fp64"]["v"].ravel()])
        err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    else:
        err = None
    main_label = "mixed" if "mixed" in results else "fp64"
    r = results[main_label]
    out["value"] = K / r["wall"]
    out["ms_per_step"] = 1e3 * r["wall"] / K
    out["dtype"] = "f32+f64" if main_label == "mixed" else "f64"
    info_last = r["infos"][-1]
    out["config"] = {"workload": f"DIVA floating channel {N}x{N} dx={512 / N:g} km, Nsigma=10, Arrhenius A, "
                                 f"MG-PCG (PointJacobi V(2,2)), Picard 1e-6, linear 1e-8; mixed: FP32 V-cycle with "
                                 f"{args.coarse} coarse operators",
                     "grid": [N, N], "precision": main_label,
                     "picard_iterations": info_last["picard_iterations"],
                     "cg_iterations": info_last["total_linear_iterations"]}
    if "fp64" in results:
        f = results["fp64"]
        out["fp64_solves_per_s"] = K / f["wall"]
        out["fp64_ms_per_solve"] = 1e3 * f["wall"] / K
        out["fp64_picard_iterations"] = f["infos"][-1]["picard_iterations"]
        out["fp64_cg_iterations"] = f["infos"][-1]["total_linear_iterations"]
    if err is not None:
        out["mixed_rel_l2_err_vs_fp64"] = err
        out["mixed_speedup_vs_fp64"] = out["value"] / out["fp64_solves_per_s"]

    # roofline: one more solve per kernel with HIP-event brackets on the solver's stream, for the three HBM-bound
    # operator/smoother kernels of the headline solve; the one with the largest device time per solve is `roofline`
    if args.no_roofline:
        print(json.dumps(out), flush=True)
        return
    m, setup = models[main_label]
    opts = setup[2]
    n0 = N * N
    n1 = (N // 2) * (N // 2)   # first coarse level (halved in both directions at these sizes)
    T = 4 if main_label == "mixed" else 8
    kinds = [("sweep", abi.PROFILE_SWEEP, f"k_sweep_m<{'float' if T == 4 else 'double'}> (finest-grid l1-Jacobi sweep)",
              n0 * (10 * T + 2), f"{N}^2 x (10 planes x {T} B: x, b, eta, etaff, beta_u, beta_v read, x written; 2 B mask)"),
             ("operator", abi.PROFILE_OPERATOR, "k_apply_mt<double> (PCG operator q = A p with p.q, p through LDS)",
              n0 * (8 * 8 + 2), f"{N}^2 x (8 planes x 8 B: p, eta, etaff, beta_u, beta_v read, q written; 2 B mask)")]
    if main_label == "mixed" and args.coarse == "galerkin":
        kinds.append(("galerkin_sweep", abi.PROFILE_GALERKIN_SWEEP,
                      f"k_gp_sweep_t<float> (Galerkin stencil sweep, symmetric packed planes, x through LDS, {N // 2}^2 level)",
                      n1 * 40 * 4, f"{N // 2}^2 x 40 planes x 4 B (32 packed couplings, 2 x 1/l1, x, b read; x written)"))
    kernels = []
    for name, kind, label, nbytes, formula in kinds:
        opts.profile_kernel = kind   # same check interval as the timed solves; eager launches (no graph) while profiling
        m.set_options(opts)
        reset_state(m)
        pinfo = m.update_velocities(geometry=True)
        launches = max(pinfo["profiled_launches"], 1)
        dur_s = pinfo["profiled_ms"] / launches / 1e3
        achieved = nbytes / dur_s / 1e9 if dur_s > 0 else 0.0
        log(f"[bench] profiled {name}: {launches} launches, avg {dur_s * 1e6:.2f} us, {achieved:.0f} GB/s "
            f"({achieved / HBM_PEAK_GBS:.2f} of peak)")
        kernels.append({"name": name, "kernel": label, "launches": launches, "avg_launch_us": dur_s * 1e6,
                        "ms_per_solve": pinfo["profiled_ms"], "algorithmic_bytes_per_launch": nbytes, "bytes_formula": formula,
                        "achieved": achieved, "frac": achieved / HBM_PEAK_GBS})
    opts.profile_kernel = abi.PROFILE_NONE
    m.set_options(opts)
    top = max(kernels, key=lambda k: k["ms_per_solve"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{top['name']}_{N}.json")
    if os.path.exists(tfile):   # measured by rocprofv3 --pmc runs of --traffic-probe, calibrated on a 1 GiB copy
        traffic = json.load(open(tfile))["traffic_bytes_per_launch"]
    out["roofline"] = {"bound": "hbm", "achieved": top["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": top["frac"],
                       "traffic": traffic, "kernel": top["kernel"], "avg_launch_us": top["avg_launch_us"],
                       "launches": top["launches"], "algorithmic_bytes_per_launch": top["algorithmic_bytes_per_launch"],
                       "kernels": kernels}
    if not args.no_cpu_baseline and "fp64" in results:
        out["cpu_baseline"] = cpu_baseline(models["fp64"][1], results["fp64"]["infos"][-1])
    if args.serve > 0:
        for m, _ in models.values():
            m.close()
        out["serving"] = serving_leg(args.serve, args.serve_rate, args.serve_slots, min(args.serve_serial, args.serve))
        out["served_p50_ms"] = out["serving"]["p50_ms"]
        out["served_p99_ms"] = out["serving"]["p99_ms"]
    if args.transport > 0:
        out["transport"] = transport_leg(N, args.transport)
        out["energy"] = energy_leg(N, 10)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
