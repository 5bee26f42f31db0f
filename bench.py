"""Benchmark: velocity solves/s at 1024² (DIVA, MG-PCG), mixed precision vs FP64, on one MI355X.

A step is one full velocity solve (update_velocities!: driving stress, Picard with DIVA coefficient updates and
multigrid-preconditioned CG, velocity-profile reconstruction) of the 1024 × 1024 floating-channel workload
(the DIVA shelf, 512 km square at dx = 500 m), from u = v = 0 and zero DIVA state, with every input already in HBM.
`value` is the mixed-precision form's solves/s (FP32 V-cycle with Galerkin coarse operators inside FP64 PCG). Beside
it: the FP64 form (the reference algorithm: rediscretized coarse levels) and FP64 with the same Galerkin coarse
operators as the mixed form, so mixed ÷ FP64-Galerkin is the precision factor alone and FP64-Galerkin ÷ FP64 the
preconditioner's; each result's relative L2 error against the FP64 solve.

Roofline: the HBM-bound kernels of the headline solve with ≥ 5 % of its device time — the PCG operator
(k_apply_mt<double>), the PCG update (k_update_xr<double>), the finest-grid l1-Jacobi sweep (k_sweep_m) and the
Galerkin stencil sweep on the first coarse level (k_gp_sweep_t) — each timed with HIP events on the solver's own
stream (nk_solver_options.profile_kernel) during one extra solve; `roofline` is the one with the most device time per
solve, the others are listed under roofline.kernels with their algorithmic bytes per launch. (The LDS-resident coarse
tail, k_g_tail, and the one-block finishers are latency-bound and have no HBM roofline.)

CPU baseline: the CPU oracle (the reference algorithm in C, OpenMP over rows) on every host thread this process may
use. `cpu_baseline` is the 1024² workload extrapolated from a bounded sample (one DIVA coefficient update and three
MG-PCG iterations, scaled with the FP64 GPU run's Picard and CG counts); `cpu_full_solves` times complete solves on all
threads and on one: the marine sheet (640 × 80, full Picard) and the reference's MISMIP+ benchmark solves at 1 km,
beside the reference's published one-thread times.

MISMIP+ scaling (`mismip_scaling`): the reference's only published benchmark (benchmarks/multigrid_scaling.jl) on the
device at 8 km to 0.25 km, iterations and times beside the published H200 ones.

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


def host_cores():
    """(threads this process may run on, CPU model) of the host."""
    n = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def cpu_baseline(setup, fp64_info, threads):
    """Oracle (C, OpenMP over rows on `threads` host threads) on a bounded sample of the 1024² workload, scaled to one
    solve with the FP64 GPU run's trip counts (extrapolated: a complete 1024² solve takes hours on the host)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)

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
    log(f"[bench] cpu oracle ({threads} threads): coefficient update {t_coef:.2f} s, solve setup {t_setup:.2f} s, CG iteration "
        f"{t_iter:.3f} s → {t_solve:.1f} s per solve ({picard} Picard, {cg} CG iterations; extrapolated)")
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"extrapolated: CPU oracle at 1024² on {threads} threads, one DIVA coefficient update ({t_coef:.2f} s), "
                      f"MG-PCG setup ({t_setup:.2f} s) and 3 MG-PCG iterations ({t_iter:.3f} s each), scaled by the FP64 "
                      f"GPU solve's {picard} Picard and {cg} CG iterations; complete solves timed in cpu_full_solves"}


def cpu_full_solves(threads_all):
    """Complete solves of the oracle (the reference algorithm in C) on the host, all visible threads and one:
    the marine sheet (640 × 80 at 1 km, FP64 MG-PCG, full Picard) and the timed solves of the reference's
    multigrid scaling benchmark at 1 km (MISMIP+, one SSA solve from zero to 1e-8 after FixedIterations(5) Picard;
    benchmarks/multigrid_scaling.jl), beside the reference's published one-thread EPYC 9745 times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = {}
    for threads in sorted({threads_all, 1}, reverse=True):
        oracle.set_threads(threads)
        g, p, o, f = workloads.marine_sheet()
        m = oracle.OracleModel(g, p, o)
        workloads.load_into(m, f)
        t0 = time.perf_counter()
        info = m.update_velocities()
        t_marine = time.perf_counter() - t0
        g, p, o, f = workloads.mismip_scaling(1000.0)
        m = oracle.OracleModel(g, p, o)
        workloads.load_into(m, f)
        m.update_velocities()
        times = {}
        for name, pre in (("jacobi", abi.PRECONDITIONER_DIAGONAL), ("multigrid", abi.PRECONDITIONER_MULTIGRID)):
            oo = abi.default_options(linear=abi.convergence(1e-8, 0.0, 100_000), picard=o.picard, preconditioner=pre)
            m.set_options(oo)
            m.set("U", 0.0)
            m.set("V", 0.0)
            t0 = time.perf_counter()
            it = m.solve()["linear_iterations"]
            times[name] = (time.perf_counter() - t0, it)
        log(f"[bench] cpu oracle, {threads} threads: marine sheet 640x80 {t_marine:.2f} s ({info['picard_iterations']} Picard, "
            f"{info['total_linear_iterations']} CG); MISMIP+ 1 km solve Jacobi {times['jacobi'][0]:.3f} s ({times['jacobi'][1]} it), "
            f"multigrid {times['multigrid'][0]:.3f} s ({times['multigrid'][1]} it)")
        out[f"threads_{threads}"] = {"marine_sheet_640x80_s": t_marine, "marine_picard": info["picard_iterations"],
                                     "marine_cg": info["total_linear_iterations"],
                                     "mismip_1km_jacobi_s": times["jacobi"][0], "mismip_1km_jacobi_it": times["jacobi"][1],
                                     "mismip_1km_multigrid_s": times["multigrid"][0],
                                     "mismip_1km_multigrid_it": times["multigrid"][1]}
    out["published_epyc9745_1thread_mismip_1km_s"] = {"jacobi": 0.22, "multigrid": 0.10}
    t1 = out["threads_1"]
    # the oracle is a checker written for clarity (every halo, mask and division as the reference states it), not a
    # tuned CPU solver: its one-thread time over the reference's published one is stated so no ratio is overread
    out["oracle_over_published_1thread"] = {"jacobi": t1["mismip_1km_jacobi_s"] / 0.22, "multigrid": t1["mismip_1km_multigrid_s"] / 0.10}
    return out


def mismip_leg(resolutions):
    """The reference's only published benchmark (benchmarks/multigrid_scaling.jl; table at docs/src/model/numerics.md:
    42-49) on the HIP path: per resolution, update_velocities! with FixedIterations(5) Picard (as set! runs it), then
    one solve from zero to 1e-8 with each preconditioner, timed on the host clock around the second of two solves (the
    reference times the second solve! too), beside the published iterations and H200 times. FP64 is the reference
    algorithm; the mixed form (FP32 Galerkin V-cycle) is listed beside it."""
    rows = []
    for dx in resolutions:
        row = {"dx_m": dx}
        for form, prec, coarse in (("fp64", abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED),
                                   ("mixed", abi.PRECISION_MIXED, abi.COARSE_GALERKIN)):
            g, p, o, f = workloads.mismip_scaling(dx, precision=prec, coarse_operator=coarse)
            m = HIPModel(g, p, o)
            workloads.load_into(m, f)
            m.update_velocities()
            pres = (("multigrid", abi.PRECONDITIONER_MULTIGRID),) if form == "mixed" else \
                (("jacobi", abi.PRECONDITIONER_DIAGONAL), ("multigrid", abi.PRECONDITIONER_MULTIGRID))
            for name, pre in pres:
                oo = abi.default_options(**{k: getattr(o, k) for k, _ in o._fields_})
                oo.linear, oo.preconditioner = abi.convergence(1e-8, 0.0, 100_000), pre
                if pre != abi.PRECONDITIONER_MULTIGRID:
                    oo.precision, oo.coarse_operator = abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED
                m.set_options(oo)
                for rep in range(2):
                    m.set("U", 0.0)
                    m.set("V", 0.0)
                    t0 = time.perf_counter()
                    info = m.solve()
                    wall = time.perf_counter() - t0
                row[f"{form}_{name}_it"] = info["linear_iterations"]
                row[f"{form}_{name}_s"] = wall
            row["grid"] = [g.Nx, g.Ny]
            m.close()
        pub_it, pub_t = workloads.MISMIP_SCALING_ITERATIONS[dx], workloads.MISMIP_SCALING_TIMES[dx]
        row["published_it"] = {"jacobi": pub_it[0], "multigrid": pub_it[1]}
        row["published_h200_s"] = {"jacobi": pub_t[2], "multigrid": pub_t[3]}
        row["published_cpu_1thread_s"] = {"jacobi": pub_t[0], "multigrid": pub_t[1]}
        log(f"[bench] MISMIP+ {row['grid'][0]}x{row['grid'][1]}: FP64 Jacobi {row['fp64_jacobi_it']} it {1e3 * row['fp64_jacobi_s']:.1f} ms, "
            f"multigrid {row['fp64_multigrid_it']} it {1e3 * row['fp64_multigrid_s']:.1f} ms (published {pub_it}, H200 "
            f"{pub_t[2]} / {pub_t[3]} s); mixed multigrid {row['mixed_multigrid_it']} it {1e3 * row['mixed_multigrid_s']:.1f} ms")
        rows.append(row)
    return rows


def serving_leg(n, rate_hz, slots, serial_sample, cpu_oracle=True):
    """Config 5 on a bounded sample: n requests of the stated stream (Poisson arrivals at rate_hz, sides log-uniform
    in [64, 1024] rounded to 16) through the server at their arrival times (latency at a stated load), then the same
    requests all at t = 0 (saturated: the server's capacity), and the first serial_sample of them solved one at a
    time (one slot, saturated) in the mixed form and in FP64 on the reference's rediscretized hierarchy."""
    from nunatak_hip.server import serve
    reqs = workloads.serving_stream(n=n, rate_hz=rate_hz, seed=5)
    log(f"[bench] serving {n} requests at {rate_hz} /s on {slots} slots")
    out = serve(reqs, slots=slots, progress=max(1, n // 10))
    log(f"[bench] served at {rate_hz} /s: p50 {out['p50_ms']:.0f} ms, p99 {out['p99_ms']:.0f} ms, {out['solves_per_s']:.2f} solves/s")
    sat = serve([(0.0,) + r[1:] for r in reqs], slots=slots)
    log(f"[bench] saturated ({n} requests at t = 0, {slots} slots): {sat['solves_per_s']:.2f} solves/s, p50 {sat['p50_ms']:.0f} ms, "
        f"p99 {sat['p99_ms']:.0f} ms")
    sample = [(0.0,) + r[1:] for r in reqs[:serial_sample]]
    serial_mixed = serve(sample, slots=1)
    sample64 = []
    for _, g, p, o, f in sample:
        o64 = abi.default_options(linear=o.linear, picard=o.picard, preconditioner=o.preconditioner,
                                  precision=abi.PRECISION_FP64, coarse_operator=abi.COARSE_REDISCRETIZED)
        sample64.append((0.0, g, p, o64, f))
    serial64 = serve(sample64, slots=1)
    sides = [r[1].Nx for r in reqs]
    log(f"[bench] serial (one slot) on the first {serial_sample}: mixed {serial_mixed['solves_per_s']:.2f} solves/s, "
        f"FP64 {serial64['solves_per_s']:.2f} solves/s")
    cpu_small = serving_cpu_small(sample64, serial64["records"], serial_mixed["records"]) if cpu_oracle else None
    return {"requests": n, "rate_hz": rate_hz, "slots": slots, "sides_min_max": [min(sides), max(sides)],
            "p50_ms": out["p50_ms"], "p99_ms": out["p99_ms"], "solves_per_s": out["solves_per_s"],
            "saturated_solves_per_s": sat["solves_per_s"], "saturated_p50_ms": sat["p50_ms"], "saturated_p99_ms": sat["p99_ms"],
            "serial_sample": serial_sample, "serial_mixed_solves_per_s": serial_mixed["solves_per_s"],
            "serial_fp64_solves_per_s": serial64["solves_per_s"], "cpu_oracle_smallest": cpu_small,
            "workload": "serving stream (seed 5): SSA grounded/floating channel per request, mixed precision "
                        "(FP32 Galerkin V-cycle), MG-PCG linear 1e-8, Picard 1e-6; serial baselines: the first "
                        f"{serial_sample} requests through one slot"}


def serving_request_leg(side, warmup, steps, graphs):
    """One request of the serving sample (the first of the 48-request schedule with this side) solved alone on one
    handle: warmup solves, then steps timed solves from zero velocity (wall time around update_velocities)."""
    _, sides = workloads.serving_schedule(48, 2.0, 5, 64, 1024)
    picks = [k for k, sd in enumerate(sides) if sd == side]
    if not picks:
        raise SystemExit(f"no request with side {side} in the serving sample (sides: {sorted(set(sides.tolist()))})")
    k = picks[0]
    g, p, o, f = workloads.serving_stream_request(k, sides)
    o.graphs = graphs
    m = HIPModel(g, p, o)
    times, info = [], None
    for rep in range(warmup + steps):
        workloads.load_into(m, f)
        m.synchronize()
        t0 = time.perf_counter()
        info = m.update_velocities()
        m.synchronize()
        if rep >= warmup:
            times.append((time.perf_counter() - t0) * 1e3)
    m.close()
    log(f"[bench] serving request {k} ({side}²): {np.median(times):.1f} ms median of {steps}, picard "
        f"{info['picard_iterations']}, cg {info['total_linear_iterations']}")
    return {"request": k, "side": side, "ms": times, "median_ms": float(np.median(times)), "graphs": graphs,
            "picard_iterations": info["picard_iterations"], "cg_iterations": info["total_linear_iterations"],
            "workload": "serving stream (seed 5) request: SSA grounded/floating channel, mixed precision (FP32 Galerkin "
                        "V-cycle), MG-PCG linear 1e-8, Picard 1e-6"}


def serving_cpu_small(sample64, rec64, rec_mixed, budget_s=20.0):
    """The reference algorithm on the host for the serving stream: the oracle (FP64, rediscretized multigrid, all host
    threads) solves the sample's requests from the smallest up until budget_s of host time is spent; the same requests'
    device times (serial, one slot) in FP64 and mixed come from the serial runs' records (start → finish)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads, cpu_model = host_cores()
    oracle.set_threads(threads)
    order = sorted(range(len(sample64)), key=lambda k: sample64[k][1].Nx * sample64[k][1].Ny)
    spent, done = 0.0, []
    for k in order:
        _, g, p, o, f = sample64[k]
        m = oracle.OracleModel(g, p, o)
        workloads.load_into(m, f)
        t0 = time.perf_counter()
        m.update_velocities()
        dt = time.perf_counter() - t0
        m.close()
        spent += dt
        done.append(k)
        if spent >= budget_s:
            break
    gpu64 = sum((rec64[k]["finish_ms"] - rec64[k]["start_ms"]) / 1e3 for k in done)
    gpumx = sum((rec_mixed[k]["finish_ms"] - rec_mixed[k]["start_ms"]) / 1e3 for k in done)
    sides = sorted({sample64[k][1].Nx for k in done})
    log(f"[bench] cpu oracle ({threads} threads) on the {len(done)} smallest serving requests (sides {sides[0]}–{sides[-1]}): "
        f"{len(done) / spent:.2f} solves/s; device serial FP64 {len(done) / gpu64:.2f}, mixed {len(done) / gpumx:.2f} solves/s")
    return {"requests": len(done), "sides": sides, "threads": threads, "cpu_model": cpu_model,
            "cpu_solves_per_s": len(done) / spent, "device_serial_fp64_solves_per_s": len(done) / gpu64,
            "device_serial_mixed_solves_per_s": len(done) / gpumx,
            "note": "oracle = the reference algorithm restated in C (FP64, rediscretized multigrid), complete solves; "
                    "device times are start → finish of the same requests in the serial runs (inputs to results on host)"}


def transport_leg(N, steps, advection=abi.ADVECTION_UPWIND1):
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
            m.compute_thickness_tendency(advection)
            m.thickness_substep(dt / beta, k == 2)

    step()
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    m.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    # WENO(order = 5, bounds): the limiter pass reads H and writes θ, and the tendency reads θ too
    bounded = advection == abi.ADVECTION_WENO5_BOUNDED
    bytes_step = 8.0 * N * N * (3 * (8 + (3 if bounded else 0)) + 3 * 5 + 2)
    name = {abi.ADVECTION_UPWIND1: "UpwindBiased(order = 1)", abi.ADVECTION_CENTERED2: "Centered(order = 2)",
            abi.ADVECTION_WENO5: "WENO(order = 5)", abi.ADVECTION_WENO5_BOUNDED: "WENO(order = 5, bounds = (0, 1e5))"}[advection]
    log(f"[bench] transport {N}x{N} {name}: {ms:.3f} ms per RK3 step, {bytes_step / (ms * 1e-3) / 1e9:.0f} GB/s (algorithmic, wall)")
    m.close()
    return {"workload": f"thickness transport {N}x{N}, (Bounded, Periodic), {name}, SplitRungeKutta3, "
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
    ap.add_argument("--no-cpu-full", action="store_true", help="skip the complete CPU solves (marine sheet, MISMIP+ 1 km)")
    ap.add_argument("--no-fp64-galerkin", action="store_true", help="skip the FP64 solve with Galerkin coarse operators")
    ap.add_argument("--refined", action="store_true",
                    help="also time the FP32-PCG-in-FP64-refinement form (≈ 50 s per 1024² solve: opt-in)")
    ap.add_argument("--mismip", default="8000,4000,2000,1000,500,250",
                    help="resolutions (m) of the reference's multigrid scaling benchmark run on the device ('' to skip)")
    ap.add_argument("--only", choices=["fp64", "fp64_galerkin", "mixed", "mixed_refined"], default=None)
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
    ap.add_argument("--serve-serial", type=int, default=48, help="requests timed serially (mixed and FP64) for the baselines")
    ap.add_argument("--transport", type=int, default=20, help="RK3 thickness steps timed in the transport leg (0: skip)")
    ap.add_argument("--transport-only", action="store_true",
                    help="run only the thickness-transport and enthalpy legs; prints their JSON objects")
    ap.add_argument("--serve-request", type=int, default=0,
                    help="time one request of the serving sample with this side alone (--warmup, then --steps solves on "
                         "one handle, mixed form; --graphs 0 for external profilers); prints one JSON object and exits")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the velocity solve of one grid is not "
              "sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    N, K, Wm = args.size, args.steps, args.warmup
    if args.transport_only:
        print(json.dumps({"transport": transport_leg(N, max(args.transport, 1)),
                          "transport_weno_bounded": transport_leg(N, max(args.transport, 1), abi.ADVECTION_WENO5_BOUNDED),
                          "energy": energy_leg(N, 10)}), flush=True)
        return
    if args.serve_request:
        print(json.dumps({"serving_request": serving_request_leg(args.serve_request, Wm, K, args.graphs)}), flush=True)
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
    setups = {}
    legs = [("fp64", abi.PRECISION_FP64), ("fp64_galerkin", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED),
            ("mixed_refined", abi.PRECISION_MIXED_REFINED)]
    for label, prec in legs:
        if args.only and args.only != label:
            continue
        if label == "fp64_galerkin" and (args.no_fp64_galerkin or args.coarse != "galerkin"):
            continue
        if label == "mixed_refined" and not (args.refined or args.only == "mixed_refined") or \
                (label == "mixed_refined" and args.coarse != "galerkin"):
            continue
        coarse = abi.COARSE_GALERKIN if (label != "fp64" and args.coarse == "galerkin") else abi.COARSE_REDISCRETIZED
        m, setup = make(N, prec, picard=args.picard, graphs=args.graphs, coarse=coarse)
        models[label] = (m, setup)
        setups[label] = setup
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
    main_label = "mixed" if "mixed" in results else next(k for k in ("fp64", "fp64_galerkin", "mixed_refined") if k in results)
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
    if "mixed_refined" in results:
        mr = results["mixed_refined"]
        out["mixed_refined_solves_per_s"] = K / mr["wall"]
        out["mixed_refined_ms_per_solve"] = 1e3 * mr["wall"] / K
        out["mixed_refined_picard_iterations"] = mr["infos"][-1]["picard_iterations"]
        out["mixed_refined_inner_cg_iterations"] = mr["infos"][-1]["total_linear_iterations"]
        if "fp64" in results:
            a = np.concatenate([mr["u"].ravel(), mr["v"].ravel()])
            b = np.concatenate([results["fp64"]["u"].ravel(), results["fp64"]["v"].ravel()])
            out["mixed_refined_rel_l2_err_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    if "fp64_galerkin" in results:
        fg = results["fp64_galerkin"]
        out["fp64_galerkin_solves_per_s"] = K / fg["wall"]
        out["fp64_galerkin_ms_per_solve"] = 1e3 * fg["wall"] / K
        out["fp64_galerkin_picard_iterations"] = fg["infos"][-1]["picard_iterations"]
        out["fp64_galerkin_cg_iterations"] = fg["infos"][-1]["total_linear_iterations"]
        if "fp64" in results:
            a = np.concatenate([fg["u"].ravel(), fg["v"].ravel()])
            b = np.concatenate([results["fp64"]["u"].ravel(), results["fp64"]["v"].ravel()])
            out["fp64_galerkin_rel_l2_err_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            out["fp64_galerkin_speedup_vs_fp64"] = out["fp64_galerkin_solves_per_s"] / out["fp64_solves_per_s"]
        if "mixed" in results:
            # the precision factor alone: the same Galerkin hierarchy and trip structure, FP32 against FP64 V-cycle
            out["mixed_speedup_vs_fp64_galerkin"] = out["value"] / out["fp64_galerkin_solves_per_s"]

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
    kinds.append(("update", abi.PROFILE_UPDATE, "k_update_xr<double> (PCG x += a p, r -= a q, r.r" +
                  ("; FP32 copy of r)" if main_label == "mixed" else ")"),
                  n0 * (12 * 8 + (8 if main_label == "mixed" else 0)),
                  f"{N}^2 x (12 x 8 B: x, p, q, r read, x, r written" + ("; r as FP32 written, 8 B)" if main_label == "mixed" else ")")))
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
        cores, cpu_model = host_cores()
        for mm, _ in models.values():
            mm.close()
        models = {}
        out["cpu_baseline"] = cpu_baseline(setups["fp64"], results["fp64"]["infos"][-1], cores)
        out["cpu_baseline"]["cpu_model"] = cpu_model
        if not args.no_cpu_full:
            out["cpu_full_solves"] = cpu_full_solves(cores)
    if args.mismip:
        out["mismip_scaling"] = mismip_leg([float(x) for x in args.mismip.split(",")])
    if args.serve > 0:
        for m, _ in models.values():
            m.close()
        models = {}
        out["serving"] = serving_leg(args.serve, args.serve_rate, args.serve_slots, min(args.serve_serial, args.serve),
                                     cpu_oracle=not args.no_cpu_baseline)
        out["served_p50_ms"] = out["serving"]["p50_ms"]
        out["served_p99_ms"] = out["serving"]["p99_ms"]
    if args.transport > 0:
        out["transport"] = transport_leg(N, args.transport)
        out["transport_weno_bounded"] = transport_leg(N, args.transport, abi.ADVECTION_WENO5_BOUNDED)
        out["energy"] = energy_leg(N, 10)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
