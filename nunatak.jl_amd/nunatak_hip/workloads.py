"""Seeded synthetic inputs of the benchmark workloads (numpy PCG64; SI units; year = 31 556 926 s).

Each generator returns (grid, physics, options, fields) where fields maps field names to arrays shaped like the
field ((ny, nx) or (nz, ny, nx)). Decisions where a workload statement has no direct counterpart in the
reference's semantics:

* Inflow at the west wall: Bounded walls are impenetrable identity rows in the reference
  (src/StressBalances/ssa_operator.jl:14-17), so the inflow face carries u = 0 and the channel is driven by its own
  thickness gradient.
* Calving front at the east end: the front force acts only on faces with ice on one side
  (src/StressBalances/driving_stress.jl:64-70), so the last column of cells is open ocean (H = 0).
* 3x3 box filter: replicate edges. Gaussian random fields: white noise filtered in Fourier space by a Gaussian of
  the stated correlation length, rescaled to the stated standard deviation.
"""
import numpy as np

from . import abi

YEAR = 31556926.0


def box3(a):
    p = np.pad(a, 1, mode="edge")
    out = np.zeros_like(a)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            out += p[1 + dj: 1 + dj + a.shape[0], 1 + di: 1 + di + a.shape[1]]
    return out / 9.0


def gaussian_random_field(rng, shape, spacing, correlation, sigma):
    ny, nx = shape
    w = rng.standard_normal(shape)
    kx = np.fft.fftfreq(nx, d=spacing[0]) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=spacing[1]) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-K2 * correlation ** 2 / 4)))
    f -= f.mean()
    s = f.std()
    return f * (sigma / s if s > 0 else 0.0)


def _centers(Nx, Ny, dx, dy):
    x = (np.arange(Nx) + 0.5) * dx
    y = (np.arange(Ny) + 0.5) * dy
    return np.meshgrid(x, y, indexing="xy")


def floating_channel(Nx=64, Ny=32, dx=1e3, seed=1, A=1e-25, scale_km=1.0, diva=False, Nsigma=10,
                     preconditioner=abi.PRECONDITIONER_DIAGONAL, linear_reltol=1e-10, picard_reltol=1e-6):
    """Config 1: floating shelf in a channel. Nx ice columns plus one open-ocean column (so the grid has Nx + 1
    cells), bed −2000 m, H = 1000 − 8·(x/km)/scale_km + box3(N(0, 10 m)), no-slip side walls, ρᵢ = 917,
    ρw = 1028, Glen n = 3 with constant A, no friction (fully afloat), Picard 1e-6, PCG 1e-10, Jacobi."""
    rng = np.random.Generator(np.random.PCG64(seed))
    NxT = Nx + 1
    g = abi.default_grid(NxT, Ny, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4, Nsigma=Nsigma)
    X, _ = _centers(NxT, Ny, dx, dx)
    H = 1000.0 - 8.0 * (X / 1e3) / scale_km + box3(rng.normal(0.0, 10.0, size=(Ny, NxT)))
    H[:, Nx:] = 0.0
    phys = abi.default_physics(ice_density=917.0, ocean_density=1028.0, grounding=abi.GROUNDING_SUBGRID,
                               walls=abi.NO_SLIP, stress_balance=abi.DIVA if diva else abi.SSA)
    opts = abi.default_options(linear=abi.convergence(linear_reltol, 0.0, 10_000),
                               picard=abi.convergence(picard_reltol, 0.0, 100), preconditioner=preconditioner)
    fields = {"H": H, "BED": np.full((Ny, NxT), -2000.0), "A": np.full((Ny, NxT), A), "C": np.zeros((Ny, NxT))}
    return g, phys, opts, fields


def marine_sheet(Nx=640, Ny=80, dx=1e3, seed=2, preconditioner=abi.PRECONDITIONER_MULTIGRID):
    """Config 2: grounded + floating marine sheet, 640 × 80 at 1 km, MG-PCG with the reference defaults.
    B = 200 − 0.9·x_km + GRF(σ = 50 m, ℓ = 10 km) + 500 m Gaussian ridges (8 km width) along both y walls;
    H = 2500·√(1 − x/600 km) for x < 600 km else 0; SubgridGroundedFraction; Weertman m = 1/3 with
    C = 3.16e6 · exp(0.3 ξ), ξ ~ N(0, 1) per cell; Glen n = 3, A = 6.3e-25."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = abi.default_grid(Nx, Ny, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4)
    X, Y = _centers(Nx, Ny, dx, dx)
    Ly = Ny * dx
    B = 200.0 - 0.9 * X / 1e3 + gaussian_random_field(rng, (Ny, Nx), (dx, dx), 10e3, 50.0)
    B += 500.0 * np.exp(-(Y / 8e3) ** 2) + 500.0 * np.exp(-((Ly - Y) / 8e3) ** 2)
    H = np.where(X < 600e3, 2500.0 * np.sqrt(np.maximum(1.0 - X / 600e3, 0.0)), 0.0)
    C = 3.16e6 * np.exp(0.3 * rng.standard_normal((Ny, Nx)))
    phys = abi.default_physics(grounding=abi.GROUNDING_SUBGRID, friction=abi.FRICTION_WEERTMAN,
                               friction_exponent=1.0 / 3.0)
    opts = abi.default_options(preconditioner=preconditioner)
    fields = {"H": H, "BED": B, "A": np.full((Ny, Nx), 6.3e-25), "C": C}
    return g, phys, opts, fields


def diva_shelf(N=1024, dx=500.0, seed=3, Nsigma=10, precision=abi.PRECISION_FP64,
               preconditioner=abi.PRECONDITIONER_MULTIGRID, linear_reltol=1e-8, picard_reltol=1e-6):
    """DIVA floating shelf: the floating channel's geometry rescaled to 512 km × 512 km (N × N cells at dx): N − 1 ice columns and
    one open-ocean column, H = 1000 − 8·(x/km)·(64 km / 512 km) + box3(N(0, 10 m)), bed −2000, no-slip walls;
    Arrhenius A on a 10-level σ grid from T(ζ) = T_s·ζ (°C), T_s ~ U[−30, 0] °C per column; regularized Coulomb
    (m = 1/3, α² = 0.5, C = 3.16e6, u₀ = 300 m/yr; the shelf is afloat so β ≡ 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = abi.default_grid(N, N, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4, Nsigma=Nsigma)
    X, _ = _centers(N, N, dx, dx)
    H = 1000.0 - 8.0 * (X / 1e3) * (64.0 / 512.0) * (512e3 / (N * dx)) + box3(rng.normal(0.0, 10.0, size=(N, N)))
    H[:, N - 1:] = 0.0
    Ts = rng.uniform(-30.0, 0.0, size=(N, N))
    zf = (np.arange(Nsigma + 1) / Nsigma) ** 2
    zc = (zf[:-1] + zf[1:]) / 2
    T = 273.15 + Ts[None, :, :] * zc[:, None, None]
    phys = abi.default_physics(ice_density=917.0, ocean_density=1028.0, grounding=abi.GROUNDING_SUBGRID,
                               walls=abi.NO_SLIP, stress_balance=abi.DIVA, rate_factor_depth=abi.RATE_FACTOR_3D,
                               friction=abi.FRICTION_REGULARIZED_COULOMB, friction_exponent=1.0 / 3.0,
                               coulomb_coefficient=0.5, velocity_regularization=300.0 / YEAR)
    opts = abi.default_options(linear=abi.convergence(linear_reltol, 0.0, 10_000),
                               picard=abi.convergence(picard_reltol, 0.0, 100), preconditioner=preconditioner,
                               precision=precision)
    fields = {"H": H, "BED": np.full((N, N), -2000.0), "C": np.full((N, N), 3.16e6), "TEMPERATURE": T}
    return g, phys, opts, fields


def friction_inversion(N=512, dx=1e3, seed=4, A=1e-25, preconditioner=abi.PRECONDITIONER_MULTIGRID,
                       linear_reltol=1e-12, picard_reltol=1e-10):
    """Friction inversion: gradient of J = ½Σ(|u| − u_obs)² with respect to log C on the floating channel's geometry at N × N, dx (512 km at
    the defaults): N − 1 ice columns and one open-ocean column, H = 1000 − 8·(x/km)·(64 km / L) + box3(N(0, 10 m)),
    no-slip walls, ρᵢ = 917, ρw = 1028, Glen n = 3 with constant A, SSA, Weertman m = 1/3.

    Config 1's bed (−2000 m) floats the whole shelf, where friction has no effect and ∂J/∂C ≡ 0, so the bed here is
    B = −300 − 1400·x/L: grounded upstream, afloat downstream, with a subgrid grounding line between.
    The prior is C₀ = 3.16e6 everywhere; the truth is log C = log C₀ + GRF(σ = 0.2, ℓ = 20 km). Returns
    (grid, physics, options, fields with C = C₀, log C_true)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = N * dx
    g = abi.default_grid(N, N, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4)
    X, _ = _centers(N, N, dx, dx)
    H = 1000.0 - 8.0 * (X / 1e3) * (64e3 / L) + box3(rng.normal(0.0, 10.0, size=(N, N)))
    H[:, N - 1:] = 0.0
    B = -300.0 - 1400.0 * X / L
    logC0 = np.full((N, N), np.log(3.16e6))
    logC_true = logC0 + gaussian_random_field(rng, (N, N), (dx, dx), 20e3, 0.2)
    phys = abi.default_physics(ice_density=917.0, ocean_density=1028.0, grounding=abi.GROUNDING_SUBGRID,
                               walls=abi.NO_SLIP, friction=abi.FRICTION_WEERTMAN, friction_exponent=1.0 / 3.0)
    opts = abi.default_options(linear=abi.convergence(linear_reltol, 0.0, 10_000),
                               picard=abi.convergence(picard_reltol, 0.0, 500), preconditioner=preconditioner)
    fields = {"H": H, "BED": B, "A": np.full((N, N), A), "C": np.exp(logC0)}
    return g, phys, opts, fields, logC_true


def serving_request(N, seed, dx=1e3, precision=abi.PRECISION_MIXED, linear_reltol=1e-8, picard_reltol=1e-6):
    """One request of the serving workload: the floating channel's geometry at N × N, dx (N − 1 ice columns and one open-ocean
    column, H = 1000 − 8·(x/km)·(64 km / L) + box3(N(0, 10 m)), no-slip walls) on the friction-inversion bed
    B = −300 − 1400·x/L (grounded upstream, afloat downstream) with its own seeded friction
    C = 3.16e6·exp(GRF(σ = 0.2, ℓ = 20 km)); SSA, Glen A = 1e-25, Weertman m = 1/3, MG-PCG at the given precision
    (the mixed form with Galerkin coarse operators), linear reltol 1e-8, Picard 1e-6."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = N * dx
    g = abi.default_grid(N, N, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4)
    X, _ = _centers(N, N, dx, dx)
    H = 1000.0 - 8.0 * (X / 1e3) * (64e3 / L) + box3(rng.normal(0.0, 10.0, size=(N, N)))
    H[:, N - 1:] = 0.0
    B = -300.0 - 1400.0 * X / L
    C = 3.16e6 * np.exp(gaussian_random_field(rng, (N, N), (dx, dx), 20e3, 0.2))
    phys = abi.default_physics(ice_density=917.0, ocean_density=1028.0, grounding=abi.GROUNDING_SUBGRID,
                               walls=abi.NO_SLIP, friction=abi.FRICTION_WEERTMAN, friction_exponent=1.0 / 3.0)
    opts = abi.default_options(linear=abi.convergence(linear_reltol, 0.0, 10_000),
                               picard=abi.convergence(picard_reltol, 0.0, 100),
                               preconditioner=abi.PRECONDITIONER_MULTIGRID, precision=precision,
                               coarse_operator=(abi.COARSE_GALERKIN if precision == abi.PRECISION_MIXED
                                                else abi.COARSE_REDISCRETIZED))
    fields = {"H": H, "BED": B, "A": np.full((N, N), 1e-25), "C": C}
    return g, phys, opts, fields


def serving_schedule(n=2000, rate_hz=4.0, seed=5, smallest=64, largest=1024):
    """(arrival_s, side) arrays of the serving stream without building its requests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = 16 * np.round(np.exp(rng.uniform(np.log(smallest), np.log(largest), size=n)) / 16).astype(int)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=n))
    return arrivals, sides


def serving_stream_request(k, sides, seed=5, precision=abi.PRECISION_MIXED):
    """Request k of the serving stream whose sides serving_schedule drew."""
    return serving_request(int(sides[k]), seed * 100003 + k, precision=precision)


def serving_stream(n=2000, rate_hz=4.0, seed=5, smallest=64, largest=1024, precision=abi.PRECISION_MIXED):
    """Config 5: n requests with Poisson arrivals at rate_hz; side = 16·round(exp(U[ln 64, ln 1024]) / 16); request
    k is serving_request(side, seed·100003 + k). Returns [(arrival_s, grid, physics, options, fields), ...]."""
    arrivals, sides = serving_schedule(n, rate_hz, seed, smallest, largest)
    out = []
    for k in range(n):
        g, p, o, f = serving_request(int(sides[k]), seed * 100003 + k, precision=precision)
        out.append((float(arrivals[k]), g, p, o, f))
    return out


# MISMIP+ marine ice stream of the reference's multigrid scaling benchmark (benchmarks/multigrid_scaling.jl:11-37)
MISMIP_LX, MISMIP_LY = 640e3, 80e3


def mismip_bed(X, Y):
    """MISMIP+ bed (Asay-Davis et al., 2016) as benchmarks/multigrid_scaling.jl:11-16 writes it."""
    B0, B2, B4, B6, xbar = -150.0, -728.8, 343.91, -50.57, 300e3
    fc, dc, wc, z_deep = 4e3, 500.0, 24e3, -720.0
    r = X / xbar
    along = B0 + B2 * r ** 2 + B4 * r ** 4 + B6 * r ** 6
    across = dc / (1 + np.exp(-2 * (Y - MISMIP_LY / 2 - wc) / fc)) + dc / (1 + np.exp(2 * (Y - MISMIP_LY / 2 + wc) / fc))
    return np.maximum(along + across, z_deep)


def mismip_scaling(dx, preconditioner=abi.PRECONDITIONER_MULTIGRID, precision=abi.PRECISION_FP64,
                   coarse_operator=abi.COARSE_REDISCRETIZED):
    """scaling_model(arch, Δx) of benchmarks/multigrid_scaling.jl:20-37: 640 km × 80 km at dx (Bounded², halo 4),
    H = 1500·√(1 − x/650 km) for x < 600 km (calving front at 600 km), MISMIP+ bed, Glen A = 2.0392e-17 /yr with
    ε̇₀ = 1e-5 /yr, regularized Coulomb C = 1e4·yr^(1/3), α² = 0.5, m = 1/3, u₀ = 1e-5 /yr, free-slip walls,
    SubgridGroundedFraction(), SmoothIceMask and the IceSheetConstants defaults; Picard FixedIterations(5) with the
    multigrid PCG at reltol 1e-6 (maxiter 50 000). set!(model, H) runs update_velocities! once with these; the timed
    solve then starts from zero with the coefficients that Picard left (benchmarks/multigrid_scaling.jl:39-50)."""
    Nx, Ny = int(round(MISMIP_LX / dx)), int(round(MISMIP_LY / dx))
    g = abi.default_grid(Nx, Ny, dx, dx, topology=(abi.BOUNDED, abi.BOUNDED), halo=4)
    X, Y = _centers(Nx, Ny, dx, dx)
    H = np.where(X < 600e3, 1500.0 * np.sqrt(np.maximum(1.0 - X / 650e3, 0.0)), 0.0)
    phys = abi.default_physics(grounding=abi.GROUNDING_SUBGRID, friction=abi.FRICTION_REGULARIZED_COULOMB,
                               friction_exponent=1.0 / 3.0, coulomb_coefficient=0.5,
                               velocity_regularization=1e-5 / YEAR, strain_rate_regularization=1e-5 / YEAR,
                               walls=abi.FREE_SLIP)
    opts = abi.default_options(linear=abi.convergence(1e-6, 0.0, 50_000), picard=abi.fixed_iterations(5),
                               preconditioner=preconditioner, precision=precision, coarse_operator=coarse_operator)
    fields = {"H": H, "BED": mismip_bed(X, Y), "A": np.full((Ny, Nx), 2.0392e-17 / YEAR),
              "C": np.full((Ny, Nx), 1e4 * YEAR ** (1.0 / 3.0))}
    return g, phys, opts, fields


# Iterations of one solve from zero to reltol 1e-8 published with the benchmark (docs/src/model/numerics.md:44-49):
# Δx (m) → (Jacobi-diagonal PCG, multigrid PCG)
MISMIP_SCALING_ITERATIONS = {8000.0: (42, 14), 4000.0: (91, 18), 2000.0: (188, 20), 1000.0: (392, 22),
                             500.0: (817, 24), 250.0: (1769, 27)}
# Published times of the same solves (s): Δx → (CPU Jacobi, CPU multigrid, H200 Jacobi, H200 multigrid); CPU is one
# thread of an AMD EPYC 9745 (docs/src/model/numerics.md:44-52)
MISMIP_SCALING_TIMES = {8000.0: (0.001, 0.002, 0.07, 0.03), 4000.0: (0.004, 0.007, 0.05, 0.08),
                        2000.0: (0.03, 0.02, 0.11, 0.04), 1000.0: (0.22, 0.10, 0.24, 0.06),
                        500.0: (1.9, 0.43, 0.59, 0.06), 250.0: (None, None, 1.3, 0.08)}


def mismip_timed_solves(model, reltol=1e-8, maxiter=100_000, jacobi=True, multigrid=True):
    """timed_solve of benchmarks/multigrid_scaling.jl:39-50 on a model (HIPModel or the oracle) created with the
    multigrid preconditioner and loaded with mismip_scaling's fields: update_velocities! (as set! runs it), then
    each preconditioner solves A x = τ from x = 0 to reltol with the coefficients Picard left. Returns
    {"jacobi": (iterations, u, v, info), "multigrid": (...)}."""
    base = model.options
    model.update_velocities()
    out = {}
    for name, pre, want in (("jacobi", abi.PRECONDITIONER_DIAGONAL, jacobi), ("multigrid", abi.PRECONDITIONER_MULTIGRID, multigrid)):
        if not want:
            continue
        o = abi.default_options(**{k: getattr(base, k) for k, _ in base._fields_})
        o.linear = abi.convergence(reltol, 0.0, maxiter)
        o.preconditioner = pre
        if pre != abi.PRECONDITIONER_MULTIGRID:
            o.precision, o.coarse_operator = abi.PRECISION_FP64, abi.COARSE_REDISCRETIZED
        model.set_options(o)
        model.set("U", 0.0)
        model.set("V", 0.0)
        info = model.solve()
        out[name] = (info["linear_iterations"], model.get("U").copy(), model.get("V").copy(), info)
    model.set_options(base)
    return out


def centre_speed(u, v):
    """|u| at cell centres of a Bounded × Bounded grid: √(ℑx(u)² + ℑy(v)²), u shaped (Ny, Nx + 1), v (Ny + 1, Nx)."""
    uc = (u[:, :-1] + u[:, 1:]) / 2
    vc = (v[:-1, :] + v[1:, :]) / 2
    return np.sqrt(uc * uc + vc * vc), uc, vc


def speed_misfit(u, v, s_obs, mask):
    """J = ½ Σ_mask (|u| − s_obs)² and its gradient (ū, v̄) with respect to the face velocities."""
    s, uc, vc = centre_speed(u, v)
    d = np.where(mask, s - s_obs, 0.0)
    J = 0.5 * float(np.sum(d * d))
    w = np.where(s > 0, d / np.where(s > 0, s, 1.0), 0.0)
    gu, gv = np.zeros_like(u), np.zeros_like(v)
    gu[:, :-1] += w * uc / 2
    gu[:, 1:] += w * uc / 2
    gv[:-1, :] += w * vc / 2
    gv[1:, :] += w * vc / 2
    return J, gu, gv


def load_into(model, fields, zero_velocity=True):
    """Copy a workload's inputs into a model (HIPModel or any object with the same protocol)."""
    for name, a in fields.items():
        model.set(name, a)
    if zero_velocity:
        model.set("U", 0.0)
        model.set("V", 0.0)
    if "TEMPERATURE" in fields:
        model.rate_factor_from_temperature()
