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

// compute_thickness_tendency! (time_stepping.jl:34-63): G = (a − melt·m) − mᵍ − ∇·(ū H); melt = 1 − φ with a
// grounding treatment, else 1; mᵍ only when coupled
__global__ __launch_bounds__(NT) void k_thickness_tendency(KG g, int adv, int grounding, int coupled, const double *__restrict__ u,
                                                           const double *__restrict__ v, const double *__restrict__ H,
                                                           const double *__restrict__ a, const double *__restrict__ m,
                                                           const double *__restrict__ phi, const double *__restrict__ mg,
                                                           double *__restrict__ G) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double u0 = ldfc(g, u, i, j), u1 = ldfc(g, u, i + 1, j);
    const double v0 = ldcf(g, v, i, j), v1 = ldcf(g, v, i, j + 1);
    const double Fx0 = g.dy * u0 * face_thickness(g, H, i, j, 0, u0, adv);
    const double Fx1 = g.dy * u1 * face_thickness(g, H, i + 1, j, 0, u1, adv);
    const double Fy0 = g.dx * v0 * face_thickness(g, H, i, j, 1, v0, adv);
    const double Fy1 = g.dx * v1 * face_thickness(g, H, i, j + 1, 1, v1, adv);
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
