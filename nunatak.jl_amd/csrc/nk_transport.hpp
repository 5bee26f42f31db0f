// nk_transport.hpp — thickness transport on the device: the tendency and the Runge-Kutta substep of the mass balance
// ∂H/∂t + ∇·(ū H) = a − m (src/Models/time_stepping.jl:15-72), one thread per cell, HBM-bound.
//
// Flux form of Oceananigans' div_Uc on a uniform rectilinear grid with a Flat vertical (unit Δz): face fluxes
// F = A·u·Ĥ with A = Δy on x-faces and Δx on y-faces, Ĥ the scheme's face thickness, and
// ∇·(ū H) = (1/(Δx Δy))·((Fx(i+1) − Fx(i)) + (Fy(j+1) − Fy(j))). Thickness beyond a Bounded edge is the edge cell's
// (zero-gradient halo); the wall faces carry u = 0, so no flux crosses them.
#pragma once
#include "nk_device.hpp"

namespace nk {

// face thickness between cells (i − 1, j) and (i, j) for x-faces (dir 0), (i, j − 1) and (i, j) for y-faces (dir 1):
// UpwindBiased(order = 1) takes the upwind cell (the right one when the velocity is not positive), Centered(order = 2)
// their mean
__device__ __forceinline__ double face_thickness(const KG &g, const double *H, int i, int j, int dir, double vel, int adv) {
    const double left = dir == 0 ? ldc(g, H, i - 1, j) : ldc(g, H, i, j - 1);
    const double right = ldc(g, H, i, j);
    return adv == 0 ? (vel > 0.0 ? left : right) : (left + right) / 2.0;
}

// ---------------------------------------------------------------------------------------------------------------
// Fifth-order WENO face thickness, plain or with the bounds-preserving limiter — the reference's default thickness
// advection, WENO(order = 5, bounds = (0, 1e5)) (src/Models/ice_sheet_model.jl:123), whose limiter is updated before
// every tendency (src/Models/time_stepping.jl:37) and read face by face as in src/Models/partial_cell_front.jl:139-151.
// The scheme lives in Oceananigans (0.117), outside the reference tree; restated from its published form:
// Jiang & Shu (1996) candidate stencils and smoothness indicators with the WENO-Z weights of Borges et al. (2008)
// (αₖ = dₖ (1 + τ / (βₖ + ε)), τ = |β₀ − β₂|, ε = 1e-8), reduced to third-order WENO and then to first-order upwind
// at faces whose stencil would leave a Bounded direction, and the Zhang & Shu (2010) scaling limiter with the
// Gauss-Lobatto weight ω̂ = 5/18 (the Courant bound the reference's test states), one θ per cell covering both
// directions: face values are rescaled as θ (ψ − c) + c toward the cell mean c so that they and the interior value
// p̃ = (c − ω̂ (ψ₋ + ψ₊)) / (1 − 2ω̂) of each direction stay inside the bounds.
// ---------------------------------------------------------------------------------------------------------------
constexpr double WENO_EPS = 1e-8;
constexpr double WENO_OMEGA = 5.0 / 18.0;

// value at the right face of cell c from cells (a, b, c, d, e), upwind side on the left
__device__ __forceinline__ double weno5z(double a, double b, double c, double d, double e) {
    const double q0 = (2.0 * a - 7.0 * b + 11.0 * c) / 6.0;
    const double q1 = (-b + 5.0 * c + 2.0 * d) / 6.0;
    const double q2 = (2.0 * c + 5.0 * d - e) / 6.0;
    const double s0 = a - 2.0 * b + c, t0 = a - 4.0 * b + 3.0 * c;
    const double s1 = b - 2.0 * c + d, t1 = b - d;
    const double s2 = c - 2.0 * d + e, t2 = 3.0 * c - 4.0 * d + e;
    const double b0 = 13.0 / 12.0 * (s0 * s0) + 0.25 * (t0 * t0);
    const double b1 = 13.0 / 12.0 * (s1 * s1) + 0.25 * (t1 * t1);
    const double b2 = 13.0 / 12.0 * (s2 * s2) + 0.25 * (t2 * t2);
    const double tau = fabs(b0 - b2);
    const double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS), r2 = tau / (b2 + WENO_EPS);
    const double a0 = 0.1 * (1.0 + r0), a1 = 0.6 * (1.0 + r1), a2 = 0.3 * (1.0 + r2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// third-order WENO-Z at the right face of cell b from cells (a, b, c)
__device__ __forceinline__ double weno3z(double a, double b, double c) {
    const double q0 = (-a + 3.0 * b) / 2.0, q1 = (b + c) / 2.0;
    const double b0 = (b - a) * (b - a), b1 = (c - b) * (c - b);
    const double tau = fabs(b0 - b1);
    const double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS);
    const double a0 = (1.0 / 3.0) * (1.0 + r0), a1 = (2.0 / 3.0) * (1.0 + r1);
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}

// order of the reconstructions at face k of a direction with N cells: 5 where both biased stencils stay inside a
// Bounded direction, else 3, else 1 (Periodic: always 5)
__device__ __forceinline__ int weno_order(int k, int N, bool bounded) {
    if (!bounded) return 5;
    if (k >= 4 && k <= N - 2) return 5;
    if (k >= 3 && k <= N - 1) return 3;
    return 1;
}

// reconstruction at face (i, j) of direction dir (x: between cells i − 1 and i; y: j − 1 and j) from the upwind cell
// below the face (left) or at it (right)
__device__ __forceinline__ double weno_face(const KG &g, const double *H, int dir, int i, int j, bool left) {
    auto C = [&](int d) { return dir == 0 ? ldc(g, H, i + d, j) : ldc(g, H, i, j + d); };
    const int order = dir == 0 ? weno_order(i, g.Nx, g.tx) : weno_order(j, g.Ny, g.ty);
    if (left) {
        if (order == 5) return weno5z(C(-3), C(-2), C(-1), C(0), C(1));
        if (order == 3) return weno3z(C(-2), C(-1), C(0));
        return C(-1);
    }
    if (order == 5) return weno5z(C(2), C(1), C(0), C(-1), C(-2));
    if (order == 3) return weno3z(C(1), C(0), C(-1));
    return C(0);
}

// |(bound − c) / (extreme − c)| where the extreme differs from c, else no limit
__device__ __forceinline__ double limit_ratio(double bound, double c, double ext) {
    const double d = ext - c;
    return d == 0.0 ? 1.0 : fabs((bound - c) / d);
}

// θ of every cell (update_bounds_preserving_limiter!)
__global__ __launch_bounds__(NT) void k_weno_limiter(KG g, const double *__restrict__ H, double lower, double upper,
                                                     double *__restrict__ theta) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double c = H[o];
    double th = 1.0;
    for (int dir = 0; dir < 2; dir++) {
        const double pm = weno_face(g, H, dir, i, j, false);                                   // ψ₋: the cell's lower face
        const double pp = dir == 0 ? weno_face(g, H, 0, i + 1, j, true) : weno_face(g, H, 1, i, j + 1, true);   // ψ₊
        const double pt = (c - WENO_OMEGA * pm - WENO_OMEGA * pp) / (1.0 - 2.0 * WENO_OMEGA);
        const double M = fmax(pt, fmax(pm, pp)), m = fmin(pt, fmin(pm, pp));
        th = fmin(th, fmin(limit_ratio(upper, c, M), limit_ratio(lower, c, m)));
    }
    theta[o] = th;
}

// face thickness of the WENO schemes: the upwind cell's reconstruction, rescaled by its θ when bounded
__device__ __forceinline__ double weno_flux_thickness(const KG &g, const double *H, const double *theta, int dir, int i, int j,
                                                      double vel) {
    const double hl = weno_face(g, H, dir, i, j, true), hr = weno_face(g, H, dir, i, j, false);
    double ul = hl, ur = hr;
    if (theta) {
        const int il = dir == 0 ? i - 1 : i, jl = dir == 0 ? j : j - 1;
        const double cl = ldc(g, H, il, jl), cr = ldc(g, H, i, j);
        ul = ldc(g, theta, il, jl) * (hl - cl) + cl;
        ur = ldc(g, theta, i, j) * (hr - cr) + cr;
    }
    // upwind_biased_product: ((u + |u|) ψᴸ + (u − |u|) ψᴿ) / 2
    return ((vel + fabs(vel)) * ul + (vel - fabs(vel)) * ur) / 2.0;
}

// compute_thickness_tendency! (time_stepping.jl:34-63): G = (a − melt·m) − mᵍ − ∇·(ū H); melt = 1 − φ with a
// grounding treatment, else 1; mᵍ only when coupled. adv 0, 1: UpwindBiased(1), Centered(2); 2: WENO(5); 3: WENO(5)
// with the bounds-preserving limiter (θ from k_weno_limiter).
__global__ __launch_bounds__(NT) void k_thickness_tendency(KG g, int adv, int grounding, int coupled, const double *__restrict__ u,
                                                           const double *__restrict__ v, const double *__restrict__ H,
                                                           const double *__restrict__ a, const double *__restrict__ m,
                                                           const double *__restrict__ phi, const double *__restrict__ mg,
                                                           double *__restrict__ G, const double *__restrict__ theta) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double u0 = ldfc(g, u, i, j), u1 = ldfc(g, u, i + 1, j);
    const double v0 = ldcf(g, v, i, j), v1 = ldcf(g, v, i, j + 1);
    double Fx0, Fx1, Fy0, Fy1;
    if (adv >= 2) {
        Fx0 = g.dy * weno_flux_thickness(g, H, theta, 0, i, j, u0);
        Fx1 = g.dy * weno_flux_thickness(g, H, theta, 0, i + 1, j, u1);
        Fy0 = g.dx * weno_flux_thickness(g, H, theta, 1, i, j, v0);
        Fy1 = g.dx * weno_flux_thickness(g, H, theta, 1, i, j + 1, v1);
    } else {
        Fx0 = g.dy * u0 * face_thickness(g, H, i, j, 0, u0, adv);
        Fx1 = g.dy * u1 * face_thickness(g, H, i + 1, j, 0, u1, adv);
        Fy0 = g.dx * v0 * face_thickness(g, H, i, j, 1, v0, adv);
        Fy1 = g.dx * v1 * face_thickness(g, H, i, j + 1, 1, v1, adv);
    }
    const double div = (1.0 / (g.dx * g.dy)) * ((Fx1 - Fx0) + (Fy1 - Fy0));
    const double melt = grounding ? (1.0 - phi[o]) * m[o] : m[o];
    const double S = (a[o] - melt) - (coupled ? mg[o] : 0.0);
    G[o] = S - div;
}

// _thickness_substep! (time_stepping.jl:65-72)
__global__ __launch_bounds__(NT) void k_thickness_substep(KG g, double dtau, int final_stage, const double *__restrict__ Hm,
                                                          const double *__restrict__ G, double *__restrict__ H,
                                                          double *__restrict__ Hc) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double Hs = Hm[o] + dtau * G[o];
    H[o] = jmax(0.0, Hs);
    Hc[o] = Hc[o] + (final_stage ? jmax(0.0, -Hs) : 0.0);
}

// per-block minimum of the cell advection timescale 1/(|u|/Δx + |v|/Δy) (u, v on the cell's west and south faces)
__global__ __launch_bounds__(NT) void k_advection_timescale(KG g, const double *__restrict__ u, const double *__restrict__ v,
                                                            double *__restrict__ block_min) {
    __shared__ double sh[NT / 64];
    NK_IJ
    double t = INFINITY;
    if (i <= g.Nx && j <= g.Ny) t = 1.0 / (fabs(ldfc(g, u, i, j)) / g.dx + fabs(ldcf(g, v, i, j)) / g.dy);
    for (int s = 32; s > 0; s >>= 1) t = fmin(t, __shfl_down(t, s, 64));
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) sh[tid >> 6] = t;
    __syncthreads();
    if (tid == 0) {
        double b = sh[0];
        for (int w = 1; w < NT / 64; w++) b = fmin(b, sh[w]);
        block_min[blockIdx.x + blockIdx.y * gridDim.x] = b;
    }
}

}  // namespace nk
