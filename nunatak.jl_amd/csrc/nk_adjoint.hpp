// nk_adjoint.hpp — transpose of the SSA Picard coefficient step, for the fixed-point adjoint of picard_solve!
// (ext/NunatakEnzymeExt.jl:113-240). With the solution x* frozen as the operand, one step maps x to
// q(x) = A(c(x)) x*, c = (η̄H, η̄Hᶠᶠ = ℑxy(η̄H), βᵘ, βᵛ) from update_stress_balance_coefficients!
// (src/StressBalances/stress_balance_coefficients.jl:64-74). These kernels form ȳ = (∂q/∂x)ᵀ q̄ and, in the parameter
// pass, C̄ = (∂q/∂C)ᵀ q̄ and Ā = (∂q/∂A)ᵀ q̄, following the forward kernels' index maps (a clamped or wrapped
// neighbour receives the adjoint of every read that landed on it, which is the adjoint of Oceananigans' halo fill).
// Contributions that several threads add to one point use FP64 atomics.
#pragma once
#include "nk_device.hpp"
#include "nk_physics.hpp"

namespace nk {

__device__ __forceinline__ void aadd(double *p, double v) { atomicAdd(p, v); }

// q̄ of an operator row, zero where the row is an identity row (inactive)
__device__ __forceinline__ double qbar_u(const KG &g, const uint8_t *act, const double *qb, int i, int j) {
    const double v = qb[off(g, fx(g, i), cy(g, j))];
    return uact(g, act, i, j) ? v : 0.0;
}
__device__ __forceinline__ double qbar_v(const KG &g, const uint8_t *act, const double *qb, int i, int j) {
    const double v = qb[off(g, cx(g, i), fy(g, j))];
    return vact(g, act, i, j) ? v : 0.0;
}

// Centres: η̄H̄ from the normal stresses Txx = 2η̄H(2ε̇xx + ε̇yy), Tyy = 2η̄H(2ε̇yy + ε̇xx) of x* (operator masks).
__global__ __launch_bounds__(NT) void k_adj_centres(KG g, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                    const double *__restrict__ xv, const double *__restrict__ qbu,
                                                    const double *__restrict__ qbv, double *__restrict__ etab) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const double exx = (iu(g, act, xu, i + 1, j) - iu(g, act, xu, i, j)) / g.dx;
    const double eyy = (iv(g, act, xv, i, j + 1) - iv(g, act, xv, i, j)) / g.dy;
    const double txxb = -g.dy * qbar_u(g, act, qbu, i, j) + g.dy * qbar_u(g, act, qbu, i + 1, j);
    const double tyyb = -g.dx * qbar_v(g, act, qbv, i, j) + g.dx * qbar_v(g, act, qbv, i, j + 1);
    etab[off(g, i, j)] = txxb * 2.0 * (2.0 * exx + eyy) + tyyb * 2.0 * (2.0 * eyy + exx);
}

// Corners: η̄Hᶠᶠ̄ = T̄xy · 2ε̇xy, then η̄H̄ += ℑxyᵀ(η̄Hᶠᶠ̄) onto the (clamped / wrapped) centres.
__global__ __launch_bounds__(NT) void k_adj_corners(KG g, const uint8_t *__restrict__ act, double wf, const double *__restrict__ xu,
                                                    const double *__restrict__ xv, const double *__restrict__ qbu,
                                                    const double *__restrict__ qbv, double *__restrict__ etab) {
    NK_IJ
    if (i > nxf(g) || j > nyf(g)) return;
    double exy = ((iu(g, act, xu, i, j) - iu(g, act, xu, i, j - 1)) / g.dy + (iv(g, act, xv, i, j) - iv(g, act, xv, i - 1, j)) / g.dx) / 2.0;
    exy = periph_ff(g, i, j) ? wf * exy : exy;
    const double txyb = g.dx * qbar_u(g, act, qbu, i, j) - g.dx * qbar_u(g, act, qbu, i, j - 1) +
                        g.dy * qbar_v(g, act, qbv, i, j) - g.dy * qbar_v(g, act, qbv, i - 1, j);
    const double ffb = txyb * 2.0 * exy / 4.0;
    if (ffb == 0.0) return;
    aadd(&etab[off(g, cx(g, i - 1), cy(g, j - 1))], ffb);
    aadd(&etab[off(g, cx(g, i), cy(g, j - 1))], ffb);
    aadd(&etab[off(g, cx(g, i - 1), cy(g, j))], ffb);
    aadd(&etab[off(g, cx(g, i), cy(g, j))], ffb);
}

// Centres: η̄H̄ → Ē (η̄H ∝ Eᵉ, e = (1 − n)/2n) → strain-rate adjoints (viscosity masks: walls only); and, in the
// parameter pass, Ā (η̄H ∝ A^(−1/n) per layer).
__global__ __launch_bounds__(NT) void k_adj_viscosity(KG g, PhysK p, Sig sg, const double *__restrict__ xu, const double *__restrict__ xv,
                                                      const double *__restrict__ eta, const double *__restrict__ es, long plane,
                                                      const double *__restrict__ A, const double *__restrict__ H,
                                                      const double *__restrict__ mu, const double *__restrict__ etab,
                                                      double *__restrict__ yu, double *__restrict__ yv, double *__restrict__ exyb,
                                                      double *__restrict__ Ab, int velocity, int parameters) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double eb = etab[o];
    if (eb == 0.0) return;
    if (parameters && Ab) {
        if (!p.A3d) {
            Ab[o] += eb * p.inv_n * eta[o] / A[o];
        } else {
            const double Hm = H[o] * (mu ? mu[o] : 1.0);
            for (int k = 0; k < p.Ns; k++) {
                const long ok = o + (long)k * plane;
                Ab[ok] += eb * Hm * sg.dz[k] * es[ok] * p.inv_n / A[ok];
            }
        }
    }
    if (!velocity) return;
    const double exx = (iu_w(g, xu, i + 1, j) - iu_w(g, xu, i, j)) / g.dx;
    const double eyy = (iv_w(g, xv, i, j + 1) - iv_w(g, xv, i, j)) / g.dy;
    auto exy = [&](int a, int c) {
        double e = ((iu_w(g, xu, a, c) - iu_w(g, xu, a, c - 1)) / g.dy + (iv_w(g, xv, a, c) - iv_w(g, xv, a - 1, c)) / g.dx) / 2.0;
        return periph_ff(g, a, c) ? p.wf * e : e;
    };
    const double e00 = exy(i, j), e10 = exy(i + 1, j), e01 = exy(i, j + 1), e11 = exy(i + 1, j + 1);
    const double E = exx * exx + eyy * eyy + exx * eyy + ((e00 * e00 + e10 * e10) / 2.0 + (e01 * e01 + e11 * e11) / 2.0) / 2.0 + p.e0 * p.e0;
    const double Eb = eb * p.exp_e * eta[o] / E;
    const double exxb = Eb * (2.0 * exx + eyy), eyyb = Eb * (2.0 * eyy + exx);
    if (!periph_fc(g, i + 1, j)) aadd(&yu[off(g, fx(g, i + 1), cy(g, j))], exxb / g.dx);
    if (!periph_fc(g, i, j)) aadd(&yu[off(g, fx(g, i), cy(g, j))], -exxb / g.dx);
    if (!periph_cf(g, i, j + 1)) aadd(&yv[off(g, cx(g, i), fy(g, j + 1))], eyyb / g.dy);
    if (!periph_cf(g, i, j)) aadd(&yv[off(g, cx(g, i), fy(g, j))], -eyyb / g.dy);
    // ∂E/∂ε̇xy(corner) = ε̇xy / 2 for each of the four corners of the centre
    aadd(&exyb[off(g, fx(g, i), fy(g, j))], Eb * e00 / 2.0);
    aadd(&exyb[off(g, fx(g, i + 1), fy(g, j))], Eb * e10 / 2.0);
    aadd(&exyb[off(g, fx(g, i), fy(g, j + 1))], Eb * e01 / 2.0);
    aadd(&exyb[off(g, fx(g, i + 1), fy(g, j + 1))], Eb * e11 / 2.0);
}

// Corners: ε̇xy adjoint → velocity adjoint (viscosity masks)
__global__ __launch_bounds__(NT) void k_adj_shear(KG g, double wf, const double *__restrict__ exyb, double *__restrict__ yu,
                                                  double *__restrict__ yv) {
    NK_IJ
    if (i > nxf(g) || j > nyf(g)) return;
    const double b = exyb[off(g, i, j)];
    if (b == 0.0) return;
    const double rb = (periph_ff(g, i, j) ? wf * b : b) / 2.0;
    if (!periph_fc(g, i, j)) aadd(&yu[off(g, fx(g, i), cy(g, j))], rb / g.dy);
    if (!periph_fc(g, i, j - 1)) aadd(&yu[off(g, fx(g, i), cy(g, j - 1))], -rb / g.dy);
    if (!periph_cf(g, i, j)) aadd(&yv[off(g, cx(g, i), fy(g, j))], rb / g.dx);
    if (!periph_cf(g, i - 1, j)) aadd(&yv[off(g, cx(g, i - 1), fy(g, j))], -rb / g.dx);
}

// d basal_drag / dU and / dC (src/Frictions/frictions.jl:304-310)
__device__ __forceinline__ void drag_derivatives(const PhysK &p, double U, double C, double Nf, double &dU, double &dC) {
    const double m = p.m;
    if (p.friction == 1) {
        dU = C * (m - 1.0) * pow(U, m - 2.0);
        dC = pow(U, m - 1.0);
        return;
    }
    const double tc = p.alpha2 * Nf;
    const double D = pow(C, 1.0 / m) * U + pow(tc, 1.0 / m);
    const double Dm = pow(D, -m), Dm1 = pow(D, -m - 1.0);
    dU = C * tc * ((m - 1.0) * pow(U, m - 2.0) * Dm - m * pow(U, m - 1.0) * Dm1 * pow(C, 1.0 / m));
    dC = pow(U, m - 1.0) * tc * Dm1 * pow(tc, 1.0 / m);
}

// Faces: β̄ = Az x* q̄ on active rows; β = φ·drag(U(u, v), ℑx C, ℑx N) → velocity adjoint and, in the parameter pass, C̄.
__global__ __launch_bounds__(NT) void k_adj_drag(KG g, PhysK p, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                 const double *__restrict__ xv, const double *__restrict__ qbu,
                                                 const double *__restrict__ qbv, const double *__restrict__ C,
                                                 const double *__restrict__ N, const double *__restrict__ phiu,
                                                 const double *__restrict__ phiv, double *__restrict__ yu, double *__restrict__ yv,
                                                 double *__restrict__ Cb, int velocity, int parameters) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    if (p.friction == 0) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy;
    const bool useN = p.friction == 2;
    const double u0 = p.u0;
    if (i <= nxf(g) && j <= g.Ny && !periph_fc(g, i, j)) {
        const double bb = qbar_u(g, act, qbu, i, j) * Az * xu[o];
        if (bb != 0.0) {
            const double vm = ((ldcf(g, xv, i - 1, j) + ldcf(g, xv, i, j)) / 2.0 + (ldcf(g, xv, i - 1, j + 1) + ldcf(g, xv, i, j + 1)) / 2.0) / 2.0;
            const double uu = xu[o];
            const double U = sqrt(uu * uu + vm * vm + u0 * u0);
            const double Cf = (ldc(g, C, i - 1, j) + ldc(g, C, i, j)) / 2.0;
            const double Nf = useN ? (ldc(g, N, i - 1, j) + ldc(g, N, i, j)) / 2.0 : 0.0;
            double dU, dC;
            drag_derivatives(p, U, Cf, Nf, dU, dC);
            const double ph = phiu[o];
            if (velocity) {
                const double Ub = bb * ph * dU;
                aadd(&yu[o], Ub * uu / U);
                const double vb = Ub * (vm / U) / 4.0;
                aadd(&yv[off(g, cx(g, i - 1), fy(g, j))], vb);
                aadd(&yv[off(g, cx(g, i), fy(g, j))], vb);
                aadd(&yv[off(g, cx(g, i - 1), fy(g, j + 1))], vb);
                aadd(&yv[off(g, cx(g, i), fy(g, j + 1))], vb);
            }
            if (parameters && Cb) {
                const double cb = bb * ph * dC / 2.0;
                aadd(&Cb[off(g, cx(g, i - 1), cy(g, j))], cb);
                aadd(&Cb[off(g, cx(g, i), cy(g, j))], cb);
            }
        }
    }
    if (i <= g.Nx && j <= nyf(g) && !periph_cf(g, i, j)) {
        const double bb = qbar_v(g, act, qbv, i, j) * Az * xv[o];
        if (bb != 0.0) {
            const double um = ((ldfc(g, xu, i, j - 1) + ldfc(g, xu, i + 1, j - 1)) / 2.0 + (ldfc(g, xu, i, j) + ldfc(g, xu, i + 1, j)) / 2.0) / 2.0;
            const double vv = xv[o];
            const double U = sqrt(vv * vv + um * um + u0 * u0);
            const double Cf = (ldc(g, C, i, j - 1) + ldc(g, C, i, j)) / 2.0;
            const double Nf = useN ? (ldc(g, N, i, j - 1) + ldc(g, N, i, j)) / 2.0 : 0.0;
            double dU, dC;
            drag_derivatives(p, U, Cf, Nf, dU, dC);
            const double ph = phiv[o];
            if (velocity) {
                const double Ub = bb * ph * dU;
                aadd(&yv[o], Ub * vv / U);
                const double ub = Ub * (um / U) / 4.0;
                aadd(&yu[off(g, fx(g, i), cy(g, j - 1))], ub);
                aadd(&yu[off(g, fx(g, i + 1), cy(g, j - 1))], ub);
                aadd(&yu[off(g, fx(g, i), cy(g, j))], ub);
                aadd(&yu[off(g, fx(g, i + 1), cy(g, j))], ub);
            }
            if (parameters && Cb) {
                const double cb = bb * ph * dC / 2.0;
                aadd(&Cb[off(g, cx(g, i), cy(g, j - 1))], cb);
                aadd(&Cb[off(g, cx(g, i), cy(g, j))], cb);
            }
        }
    }
}

// q̄ = −λ over :xyz (zero elsewhere is kept by the caller's fills)
__global__ __launch_bounds__(NT) void k_negate(KG g, const double *__restrict__ au, const double *__restrict__ av, double *__restrict__ bu,
                                               double *__restrict__ bv) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    bu[o] = -au[o];
    bv[o] = -av[o];
}

// r = 1·x̄ + 1·ȳ over :xyz (copyto! then axpy!(1, ȳ, r))
__global__ __launch_bounds__(NT) void k_add(KG g, const double *__restrict__ au, const double *__restrict__ av, const double *__restrict__ bu,
                                            const double *__restrict__ bv, double *__restrict__ cu, double *__restrict__ cv) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    cu[o] = au[o] + 1.0 * bu[o];
    cv[o] = av[o] + 1.0 * bv[o];
}

}  // namespace nk
