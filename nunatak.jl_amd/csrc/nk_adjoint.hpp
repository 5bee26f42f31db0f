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

// Faces: β̄ = Az x* q̄ on active rows (or the given β̄ — the DIVA chain passes βᵇ̄); β = φ·drag(U(u, v), ℑx C, ℑx N)
// → velocity adjoint and, in the parameter pass, C̄.
__global__ __launch_bounds__(NT) void k_adj_drag(KG g, PhysK p, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                 const double *__restrict__ xv, const double *__restrict__ qbu,
                                                 const double *__restrict__ qbv, const double *__restrict__ C,
                                                 const double *__restrict__ N, const double *__restrict__ phiu,
                                                 const double *__restrict__ phiv, double *__restrict__ yu, double *__restrict__ yv,
                                                 double *__restrict__ Cb, int velocity, int parameters,
                                                 const double *__restrict__ bbar_u, const double *__restrict__ bbar_v) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    if (p.friction == 0) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy;
    const bool useN = p.friction == 2;
    const double u0 = p.u0;
    if (i <= nxf(g) && j <= g.Ny && !periph_fc(g, i, j)) {
        const double bb = bbar_u ? bbar_u[o] : qbar_u(g, act, qbu, i, j) * Az * xu[o];
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
        const double bb = bbar_v ? bbar_v[o] : qbar_v(g, act, qbv, i, j) * Az * xv[o];
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

// ---------------------------------------------------------------------------------------------------------------
// DIVA (update_diva_coefficients!, src/StressBalances/diva.jl:60-75). One step reads the lagged state
// s = (βᵇ, F₂, ησ) and writes its successor: uᵇ = u/(1 + βᵇ ℑx F₂) → βᵇ′ = φ·drag(uᵇ) → ησ′(u, βᵇ′, uᵇ, ησ) →
// η̄H, F₂′ → β = βᵇ′/(1 + βᵇ′ ℑx F₂′). The kernels below run that chain backwards; the shadows of s′ come in, those
// of s go out (the fixed-point sweep of the reference rule).
// ---------------------------------------------------------------------------------------------------------------

// Faces: β̄ = q̄·Az·x* (operator rows) through β = βᵇ/(1 + βᵇ F̄ˣ): βᵇ̄ += β̄/den², F₂̄ += −β̄ βᵇ²/den² / 2 to the two
// centres (diva.jl:35-51).
__global__ __launch_bounds__(NT) void k_adj_effective_drag(KG g, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                           const double *__restrict__ xv, const double *__restrict__ qbu,
                                                           const double *__restrict__ qbv, const double *__restrict__ bbu,
                                                           const double *__restrict__ bbv, const double *__restrict__ F2,
                                                           double *__restrict__ bbu_b, double *__restrict__ bbv_b,
                                                           double *__restrict__ F2_b) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy;
    if (i <= nxf(g) && j <= g.Ny) {
        const double bb = qbar_u(g, act, qbu, i, j) * Az * xu[o];
        if (bb != 0.0) {
            const double b = bbu[o];
            const double den = 1.0 + b * ((ldc(g, F2, i - 1, j) + ldc(g, F2, i, j)) / 2.0);
            bbu_b[o] += bb / (den * den);
            const double fb = -bb * b * b / (den * den) / 2.0;
            aadd(&F2_b[off(g, cx(g, i - 1), cy(g, j))], fb);
            aadd(&F2_b[off(g, cx(g, i), cy(g, j))], fb);
        }
    }
    if (i <= g.Nx && j <= nyf(g)) {
        const double bb = qbar_v(g, act, qbv, i, j) * Az * xv[o];
        if (bb != 0.0) {
            const double b = bbv[o];
            const double den = 1.0 + b * ((ldc(g, F2, i, j - 1) + ldc(g, F2, i, j)) / 2.0);
            bbv_b[o] += bb / (den * den);
            const double fb = -bb * b * b / (den * den) / 2.0;
            aadd(&F2_b[off(g, cx(g, i), cy(g, j - 1))], fb);
            aadd(&F2_b[off(g, cx(g, i), cy(g, j))], fb);
        }
    }
}

// Centres: DIVA viscosity and column integrals backwards (glen_flow_law.jl:122-143, column_integrals.jl). Inputs:
// η̄H̄ (etab), F̄₂′ (F2_b), ησ̄′ (es_b, σ layers). Outputs: strain-rate adjoints onto u, v and the corners (as the SSA
// kernel), βᵇ̄ and uᵇ̄ of the faces i, i + 1 (j, j + 1) through the basal shear, ησ̄ of the lagged layers (es_b_old),
// and Ā in the parameter pass.
__global__ __launch_bounds__(NT) void k_adj_viscosity_diva(KG g, PhysK p, Sig sg, const double *__restrict__ xu,
                                                           const double *__restrict__ xv, const double *__restrict__ es, long plane,
                                                           const double *__restrict__ A, const double *__restrict__ H,
                                                           const double *__restrict__ mu, const double *__restrict__ bbu,
                                                           const double *__restrict__ bbv, const double *__restrict__ ub,
                                                           const double *__restrict__ vb, const double *__restrict__ etab,
                                                           const double *__restrict__ F2_b, const double *__restrict__ es_b,
                                                           double *__restrict__ yu, double *__restrict__ yv, double *__restrict__ exyb,
                                                           double *__restrict__ bbu_b, double *__restrict__ bbv_b,
                                                           double *__restrict__ ub_b, double *__restrict__ vb_b,
                                                           double *__restrict__ es_b_old, double *__restrict__ Ab, int velocity,
                                                           int parameters) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double Hm = H[o] * (mu ? mu[o] : 1.0);
    const double eb = etab[o], fb = F2_b[o];
    // forward quantities at the fixed point
    const double exx = (iu_w(g, xu, i + 1, j) - iu_w(g, xu, i, j)) / g.dx;
    const double eyy = (iv_w(g, xv, i, j + 1) - iv_w(g, xv, i, j)) / g.dy;
    auto exy = [&](int a, int c) {
        double e = ((iu_w(g, xu, a, c) - iu_w(g, xu, a, c - 1)) / g.dy + (iv_w(g, xv, a, c) - iv_w(g, xv, a - 1, c)) / g.dx) / 2.0;
        return periph_ff(g, a, c) ? p.wf * e : e;
    };
    const double e00 = exy(i, j), e10 = exy(i + 1, j), e01 = exy(i, j + 1), e11 = exy(i + 1, j + 1);
    const double e2m = exx * exx + eyy * eyy + exx * eyy + ((e00 * e00 + e10 * e10) / 2.0 + (e01 * e01 + e11 * e11) / 2.0) / 2.0;
    const double tx = (ldfc(g, bbu, i, j) * ldfc(g, ub, i, j) + ldfc(g, bbu, i + 1, j) * ldfc(g, ub, i + 1, j)) / 2.0;
    const double ty = (ldcf(g, bbv, i, j) * ldcf(g, vb, i, j) + ldcf(g, bbv, i, j + 1) * ldcf(g, vb, i, j + 1)) / 2.0;
    double e2m_b = 0.0, tx_b = 0.0, ty_b = 0.0;
    for (int k = 0; k < p.Ns; k++) {
        const long ok = o + (long)k * plane;
        const double zeta = sg.zc[k];
        const double es0 = es[ok];
        const double em = es0 == 0.0 ? INFINITY : es0;
        const double exz = tx * (1.0 - zeta) / (2.0 * em), eyz = ty * (1.0 - zeta) / (2.0 * em);
        const double E = e2m + exz * exz + eyz * eyz + p.e0 * p.e0;
        const double pa = pow(A[p.A3d ? ok : o], p.inv_n);
        const double ek = pa * pow(E, p.exp_e) / 2.0;
        // ek feeds ησ′, η̄H = H μ Σ ek dz and F₂′ = H Σ w₂/ek
        const double ekb = es_b[ok] + eb * Hm * sg.dz[k] - fb * H[o] * sg.w2[k] / (ek * ek);
        if (parameters && Ab) {
            const double ab = ekb * ek * p.inv_n / A[p.A3d ? ok : o];
            if (p.A3d) Ab[ok] += ab;
            else Ab[o] += ab;
        }
        if (!velocity) continue;
        const double Eb = ekb * ek * p.exp_e / E;
        e2m_b += Eb;
        const double exzb = Eb * 2.0 * exz, eyzb = Eb * 2.0 * eyz;
        tx_b += exzb * (1.0 - zeta) / (2.0 * em);
        ty_b += eyzb * (1.0 - zeta) / (2.0 * em);
        // lagged layer viscosity: ∂exz/∂em = −exz/em (no dependence where the lag was still 0)
        es_b_old[ok] = es0 == 0.0 ? 0.0 : -(exzb * exz + eyzb * eyz) / em;
    }
    if (!velocity) return;
    // strain rates (as k_adj_viscosity with Ē = ∂J/∂ε̇ₑ²)
    const double exxb = e2m_b * (2.0 * exx + eyy), eyyb = e2m_b * (2.0 * eyy + exx);
    if (!periph_fc(g, i + 1, j)) aadd(&yu[off(g, fx(g, i + 1), cy(g, j))], exxb / g.dx);
    if (!periph_fc(g, i, j)) aadd(&yu[off(g, fx(g, i), cy(g, j))], -exxb / g.dx);
    if (!periph_cf(g, i, j + 1)) aadd(&yv[off(g, cx(g, i), fy(g, j + 1))], eyyb / g.dy);
    if (!periph_cf(g, i, j)) aadd(&yv[off(g, cx(g, i), fy(g, j))], -eyyb / g.dy);
    aadd(&exyb[off(g, fx(g, i), fy(g, j))], e2m_b * e00 / 2.0);
    aadd(&exyb[off(g, fx(g, i + 1), fy(g, j))], e2m_b * e10 / 2.0);
    aadd(&exyb[off(g, fx(g, i), fy(g, j + 1))], e2m_b * e01 / 2.0);
    aadd(&exyb[off(g, fx(g, i + 1), fy(g, j + 1))], e2m_b * e11 / 2.0);
    // basal shear: tx = (βᵇ(i) uᵇ(i) + βᵇ(i+1) uᵇ(i+1)) / 2
    const long ou0 = off(g, fx(g, i), cy(g, j)), ou1 = off(g, fx(g, i + 1), cy(g, j));
    const long ov0 = off(g, cx(g, i), fy(g, j)), ov1 = off(g, cx(g, i), fy(g, j + 1));
    aadd(&bbu_b[ou0], tx_b * ub[ou0] / 2.0);
    aadd(&bbu_b[ou1], tx_b * ub[ou1] / 2.0);
    aadd(&ub_b[ou0], tx_b * bbu[ou0] / 2.0);
    aadd(&ub_b[ou1], tx_b * bbu[ou1] / 2.0);
    aadd(&bbv_b[ov0], ty_b * vb[ov0] / 2.0);
    aadd(&bbv_b[ov1], ty_b * vb[ov1] / 2.0);
    aadd(&vb_b[ov0], ty_b * bbv[ov0] / 2.0);
    aadd(&vb_b[ov1], ty_b * bbv[ov1] / 2.0);
}

// Faces: uᵇ = u / (1 + βᵇ ℑx F₂) backwards (diva.jl:1-27): ū += uᵇ̄/D; the lagged βᵇ and F₂ get −uᵇ̄ uᵇ/D times
// the other factor (F₂ halved onto the two centres).
__global__ __launch_bounds__(NT) void k_adj_basal_velocity(KG g, const double *__restrict__ ub, const double *__restrict__ vb,
                                                           const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                           const double *__restrict__ F2, const double *__restrict__ ub_b,
                                                           const double *__restrict__ vb_b, double *__restrict__ yu,
                                                           double *__restrict__ yv, double *__restrict__ bbu_b_old,
                                                           double *__restrict__ bbv_b_old, double *__restrict__ F2_b_old) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    if (i <= nxf(g) && j <= g.Ny) {
        const double a = ub_b[o];
        if (a != 0.0) {
            const double F2x = (ldc(g, F2, i - 1, j) + ldc(g, F2, i, j)) / 2.0;
            const double D = 1.0 + bbu[o] * F2x;
            aadd(&yu[o], a / D);
            const double Db = -a * ub[o] / D;
            bbu_b_old[o] += Db * F2x;
            const double fb = Db * bbu[o] / 2.0;
            aadd(&F2_b_old[off(g, cx(g, i - 1), cy(g, j))], fb);
            aadd(&F2_b_old[off(g, cx(g, i), cy(g, j))], fb);
        }
    }
    if (i <= g.Nx && j <= nyf(g)) {
        const double a = vb_b[o];
        if (a != 0.0) {
            const double F2y = (ldc(g, F2, i, j - 1) + ldc(g, F2, i, j)) / 2.0;
            const double D = 1.0 + bbv[o] * F2y;
            aadd(&yv[o], a / D);
            const double Db = -a * vb[o] / D;
            bbv_b_old[o] += Db * F2y;
            const double fb = Db * bbv[o] / 2.0;
            aadd(&F2_b_old[off(g, cx(g, i), cy(g, j - 1))], fb);
            aadd(&F2_b_old[off(g, cx(g, i), cy(g, j))], fb);
        }
    }
}

// Faces: compute_velocity_profile! backwards (diva.jl:89-109): uσₖ = uᵇ (1 + βᵇ F₁ₖ), F₁ₖ = ℑx SIₖ,
// SIₖ(c) = H(c) Σ_{k′≤k} pw[k][k′] / ησ(c, k′). From ūσ: uᵇ̄ += Σ ūσₖ (1 + βᵇ F₁ₖ), βᵇ̄ += Σ ūσₖ uᵇ F₁ₖ, and
// ησ̄(c, k′) −= ūσₖ uᵇ βᵇ / 2 · H pw[k][k′] / ησ(c, k′)² for both centres c of the face.
__global__ __launch_bounds__(NT) void k_adj_velocity_profile(KG g, Sig sg, int Ns, long plane, const double *__restrict__ es,
                                                             const double *__restrict__ H, const double *__restrict__ ub,
                                                             const double *__restrict__ vb, const double *__restrict__ bbu,
                                                             const double *__restrict__ bbv, const double *__restrict__ usb,
                                                             const double *__restrict__ vsb, double *__restrict__ ub_b,
                                                             double *__restrict__ vb_b, double *__restrict__ bbu_b,
                                                             double *__restrict__ bbv_b, double *__restrict__ es_b) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    auto SI = [&](long oc, int k) {
        double I = 0.0;
        for (int kp = 0; kp <= k; kp++) I += 1.0 / es[oc + (long)kp * plane] * sg.pw[k * Ns + kp];
        return I * H[oc];
    };
    auto face = [&](double u, double b, const double *sb, long c0, long c1, double *u_b, double *b_b) {
        double ubar = 0.0, bbar = 0.0;
        for (int k = 0; k < Ns; k++) {
            const double a = sb[o + (long)k * plane];
            if (a == 0.0) continue;
            const double F1 = (SI(c0, k) + SI(c1, k)) / 2.0;
            ubar += a * (1.0 + b * F1);
            bbar += a * u * F1;
            const double sib = a * u * b / 2.0;
            for (int kp = 0; kp <= k; kp++) {
                const double e0 = es[c0 + (long)kp * plane], e1 = es[c1 + (long)kp * plane];
                aadd(&es_b[c0 + (long)kp * plane], -sib * H[c0] * sg.pw[k * Ns + kp] / (e0 * e0));
                aadd(&es_b[c1 + (long)kp * plane], -sib * H[c1] * sg.pw[k * Ns + kp] / (e1 * e1));
            }
        }
        u_b[o] += ubar;
        b_b[o] += bbar;
    };
    if (i <= nxf(g) && j <= g.Ny)
        face(ub[o], bbu[o], usb, off(g, cx(g, i - 1), cy(g, j)), off(g, cx(g, i), cy(g, j)), ub_b, bbu_b);
    if (i <= g.Nx && j <= nyf(g))
        face(vb[o], bbv[o], vsb, off(g, cx(g, i), cy(g, j - 1)), off(g, cx(g, i), cy(g, j)), vb_b, bbv_b);
}

// b = a + c over the (Nx + 1) × (Ny + 1) footprint of one plane
__global__ __launch_bounds__(NT) void k_add_plane(KG g, const double *__restrict__ a, const double *__restrict__ c, double *__restrict__ b) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    b[o] = a[o] + c[o];
}

// one plane over the (Nx + 1) × (Ny + 1) footprint
__global__ __launch_bounds__(NT) void k_copy_plane(KG g, const double *__restrict__ a, double *__restrict__ b) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    b[o] = a[o];
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
