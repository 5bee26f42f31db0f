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


def serving_stream(n=2000, rate_hz=4.0, seed=5, smallest=64, largest=1024, precision=abi.PRECISION_MIXED):
    """Config 5: n requests with Poisson arrivals at rate_hz; side = 16·round(exp(U[ln 64, ln 1024]) / 16); request
    k is serving_request(side, seed·100003 + k). Returns [(arrival_s, grid, physics, options, fields), ...]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = 16 * np.round(np.exp(rng.uniform(np.log(smallest), np.log(largest), size=n)) / 16).astype(int)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=n))
    out = []
    for k in range(n):
        g, p, o, f = serving_request(int(sides[k]), seed * 100003 + k, precision=precision)
        out.append((float(arrivals[k]), g, p, o, f))
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
