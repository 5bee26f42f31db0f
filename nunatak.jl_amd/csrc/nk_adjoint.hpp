// nk_adjoint.hpp — transpose of the SSA Picard coefficient step, for the fixed-point adjoint of picard_solve!
// (ext/NunatakEnzymeExt.jl:113-240). With the solution x* frozen as the operand, one step maps x to
// q(x) = A(c(x)) x*, c = (η̄H, η̄Hᶠᶠ = ℑxy(η̄H), βᵘ, βᵛ) from update_stress_balance_coefficients!
// (src/StressBalances/stress_balance_coefficients.jl:10-20). These kernels form ȳ = (∂q/∂x)ᵀ q̄ and, in the parameter
// pass, C̄ = (∂q/∂C)ᵀ q̄ and Ā = (∂q/∂A)ᵀ q̄, following the forward kernels' index maps (a clamped or wrapped
// neighbour receives the adjoint of every read that landed on it, which is the adjoint of Oceananigans' halo fill).
//
// Scatter without atomics. A scatter kernel's thread (i, j) adds into points (i + a, j + b) with |a|, |b| ≤ 1 after
// the forward kernel's clamp or wrap. It runs as one launch per colour class: the points of a class lie at least
// three apart along each direction (cyclically on a Periodic one), so no two threads of a launch touch one address,
// every add is a plain load-add-store, and each sum is formed in the same order on every run (bit-reproducible
// gradients). The host enumerates the classes (adjoint_classes in nk_solver.hip).
#pragma once
#include "nk_device.hpp"
#include "nk_physics.hpp"

namespace nk {

// A colour class: i = ox + sx·t for t < nx, j = oy + sy·u for u < ny.
struct Colour {
    int ox, sx, nx, oy, sy, ny;
};
#define NK_IJ_COL(col)                                                   \
    int bx_, by_;                                                        \
    xcd_tile(bx_, by_);                                                  \
    const int ti_ = bx_ * BX + threadIdx.x, tj_ = by_ * BY + threadIdx.y; \
    if (ti_ >= col.nx || tj_ >= col.ny) return;                          \
    const int i = col.ox + col.sx * ti_, j = col.oy + col.sy * tj_;

// Plain add: the launch's colour class guarantees that no other thread of the launch writes *p.
__device__ __forceinline__ void aadd(double *p, double v) { *p += v; }

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
__global__ __launch_bounds__(NT) void k_adj_corners(KG g, Colour col, const uint8_t *__restrict__ act, double wf, const double *__restrict__ xu,
                                                    const double *__restrict__ xv, const double *__restrict__ qbu,
                                                    const double *__restrict__ qbv, double *__restrict__ etab) {
    NK_IJ_COL(col)
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
__global__ __launch_bounds__(NT) void k_adj_viscosity(KG g, Colour col, PhysK p, Sig sg, const double *__restrict__ xu, const double *__restrict__ xv,
                                                      const double *__restrict__ eta, const double *__restrict__ es, long plane,
                                                      const double *__restrict__ A, const double *__restrict__ H,
                                                      const double *__restrict__ mu, const double *__restrict__ etab,
                                                      double *__restrict__ yu, double *__restrict__ yv, double *__restrict__ exyb,
                                                      double *__restrict__ Ab, int velocity, int parameters) {
    NK_IJ_COL(col)
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
__global__ __launch_bounds__(NT) void k_adj_shear(KG g, Colour col, double wf, const double *__restrict__ exyb, double *__restrict__ yu,
                                                  double *__restrict__ yv) {
    NK_IJ_COL(col)
    if (i > nxf(g) || j > nyf(g)) return;
    const double b = exyb[off(g, i, j)];
    if (b == 0.0) return;
    const double rb = (periph_ff(g, i, j) ? wf * b : b) / 2.0;
    if (!periph_fc(g, i, j)) aadd(&yu[off(g, fx(g, i), cy(g, j))], rb / g.dy);
    if (!periph_fc(g, i, j - 1)) aadd(&yu[off(g, fx(g, i), cy(g, j - 1))], -rb / g.dy);
    if (!periph_cf(g, i, j)) aadd(&yv[off(g, cx(g, i), fy(g, j))], rb / g.dx);
    if (!periph_cf(g, i - 1, j)) aadd(&yv[off(g, cx(g, i - 1), fy(g, j))], -rb / g.dx);
}

// d basal_drag / dU and / dC (src/Frictions/frictions.jl:83-89)
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
// → velocity adjoint and, in the parameter pass, C̄. (xu, xv) is the point the drag is linearized at, (xsu, xsv) the
// vector the operator multiplies (both x* at the fixed point; x_{k−1} and x_k in an unrolled Picard step).
__global__ __launch_bounds__(NT) void k_adj_drag(KG g, Colour col, PhysK p, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                 const double *__restrict__ xv, const double *__restrict__ qbu,
                                                 const double *__restrict__ qbv, const double *__restrict__ C,
                                                 const double *__restrict__ N, const double *__restrict__ phiu,
                                                 const double *__restrict__ phiv, double *__restrict__ yu, double *__restrict__ yv,
                                                 double *__restrict__ Cb, int velocity, int parameters,
                                                 const double *__restrict__ bbar_u, const double *__restrict__ bbar_v,
                                                 const double *__restrict__ xsu, const double *__restrict__ xsv) {
    NK_IJ_COL(col)
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    if (p.friction == 0) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy;
    const bool useN = p.friction == 2;
    const double u0 = p.u0;
    if (i <= nxf(g) && j <= g.Ny && !periph_fc(g, i, j)) {
        const double bb = bbar_u ? bbar_u[o] : qbar_u(g, act, qbu, i, j) * Az * xsu[o];
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
        const double bb = bbar_v ? bbar_v[o] : qbar_v(g, act, qbv, i, j) * Az * xsv[o];
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
__global__ __launch_bounds__(NT) void k_adj_effective_drag(KG g, Colour col, const uint8_t *__restrict__ act, const double *__restrict__ xu,
                                                           const double *__restrict__ xv, const double *__restrict__ qbu,
                                                           const double *__restrict__ qbv, const double *__restrict__ bbu,
                                                           const double *__restrict__ bbv, const double *__restrict__ F2,
                                                           double *__restrict__ bbu_b, double *__restrict__ bbv_b,
                                                           double *__restrict__ F2_b) {
    NK_IJ_COL(col)
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
__global__ __launch_bounds__(NT) void k_adj_viscosity_diva(KG g, Colour col, PhysK p, Sig sg, const double *__restrict__ xu,
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
    NK_IJ_COL(col)
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
__global__ __launch_bounds__(NT) void k_adj_basal_velocity(KG g, Colour col, const double *__restrict__ ub, const double *__restrict__ vb,
                                                           const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                           const double *__restrict__ F2, const double *__restrict__ ub_b,
                                                           const double *__restrict__ vb_b, double *__restrict__ yu,
                                                           double *__restrict__ yv, double *__restrict__ bbu_b_old,
                                                           double *__restrict__ bbv_b_old, double *__restrict__ F2_b_old) {
    NK_IJ_COL(col)
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
__global__ __launch_bounds__(NT) void k_adj_velocity_profile(KG g, Colour col, Sig sg, int Ns, long plane, const double *__restrict__ es,
                                                             const double *__restrict__ H, const double *__restrict__ ub,
                                                             const double *__restrict__ vb, const double *__restrict__ bbu,
                                                             const double *__restrict__ bbv, const double *__restrict__ usb,
                                                             const double *__restrict__ vsb, double *__restrict__ ub_b,
                                                             double *__restrict__ vb_b, double *__restrict__ bbu_b,
                                                             double *__restrict__ bbv_b, double *__restrict__ es_b) {
    NK_IJ_COL(col)
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

// ---------------------------------------------------------------------------------------------------------------
// Geometry pullback of the converged solve: ∂J/∂H and ∂J/∂b. H and b enter the velocity solve through the geometry
// diagnostics (s, μ, φᵘ, φᵛ, N of update_geometry!, src/Models/ice_sheet_model_methods.jl:103-111), the driving
// stress with its calving-front and grounding-line terms (src/StressBalances/driving_stress.jl:56-125), the H μ
// factor of η̄H and the H factor of F₂ (src/StressBalances/column_integrals.jl:70-71) and the drag (φ and N,
// src/Frictions/frictions.jl:83-89). With τ̄ = λ, the coefficient shadows of the final pullback (η̄H̄ after the
// corner interpolation, F̄₂ for DIVA) and β̄ = Az x* q̄ (or the DIVA basal β̄ᵇ), the pullback runs in two passes
// without atomics: every face writes what it sends to its two cells (the cells its forward reads landed on, clamped
// or wrapped), then every cell adds the sends of the faces that read it, in a fixed order.
// ---------------------------------------------------------------------------------------------------------------

// front_force(H, s) = g (ρᵢ H² − ρw D²)/2, D = clamp(H − s, 0, H): derivatives in H and s (one-sided at the kinks)
__device__ __forceinline__ void front_force_d(double H, double s, double ri, double rw, double gr, double &dH, double &ds) {
    const double d = H - s;
    double D = d, dDdH = 1.0, dDds = -1.0;
    if (d > H) { D = H; dDdH = 1.0; dDds = 0.0; }
    else if (d < 0.0) { D = 0.0; dDdH = 0.0; dDds = 0.0; }
    dH = gr * (ri * H - rw * D * dDdH);
    ds = gr * (-rw * D * dDds);
}

// grounding_line_stress (driving_stress.jl:92-103) derivatives in (H₁, H₂, s₁, s₂, b₁, b₂)
__device__ __forceinline__ void gl_stress_d(double H1, double H2, double s1, double s2, double b1, double b2, double ri, double rw,
                                            double gr, double d[6]) {
    const double f1 = H1 + rw / ri * b1, f2 = H2 + rw / ri * b2;
    const double den = f1 - f2;
    const double theta = f1 / den;
    const double Hg = H1 + theta * (H2 - H1);
    const double sg = (1.0 - ri / rw) * Hg;
    const double c = -ri * gr / 2.0, kap = 1.0 - ri / rw;
    const double dHg = c * ((s2 - s1) + kap * (H1 - H2));
    const double t1 = -f2 / (den * den), t2 = f1 / (den * den);   // ∂θ/∂f₁, ∂θ/∂f₂
    const double dH = H2 - H1;
    d[0] = c * (sg - s1) + dHg * (1.0 - theta + dH * t1);
    d[1] = c * (s2 - sg) + dHg * (theta + dH * t2);
    d[2] = -c * (H1 + Hg);
    d[3] = c * (Hg + H2);
    d[4] = dHg * dH * t1 * (rw / ri);
    d[5] = dHg * dH * t2 * (rw / ri);
}

// segment_fraction(f₁, f₂) (subgrid_grounded_fraction.jl:31-35) derivatives
__device__ __forceinline__ void segment_fraction_d(double f1, double f2, double &d1, double &d2) {
    const double fp = jmax(f1, 0.0) + jmax(f2, 0.0), sf = fabs(f1) + fabs(f2);
    if (sf == 0.0) { d1 = d2 = 0.0; return; }
    const double p1 = signbit(f1 - 0.0) ? 0.0 : 1.0, p2 = signbit(f2 - 0.0) ? 0.0 : 1.0;
    const double g1 = f1 < 0.0 ? -1.0 : 1.0, g2 = f2 < 0.0 ? -1.0 : 1.0;
    d1 = (p1 * sf - fp * g1) / (sf * sf);
    d2 = (p2 * sf - fp * g2) / (sf * sf);
}

// ∂s/∂H, ∂s/∂b and ∂N/∂H, ∂N/∂b at a cell (update_geometry!: surface elevation and effective pressure)
struct GeoD { double sH, sb, NH, Nb; };
__device__ __forceinline__ GeoD geometry_d(const PhysK &p, double H, double b) {
    GeoD r;
    if (p.grounding != 0) {
        const bool grounded = !signbit((b + H) - (1.0 - p.ri_rw) * H);   // jmax keeps its first argument
        r.sH = grounded ? 1.0 : 1.0 - p.ri_rw;
        r.sb = grounded ? 1.0 : 0.0;
    } else {
        r.sH = 1.0; r.sb = 1.0;
    }
    const bool bneg = signbit(b - 0.0);                                   // jmin(b, 0) keeps b
    const double P = p.rig * H + p.rwg * (bneg ? b : 0.0);
    const bool pos = !signbit(P - 0.0);
    r.NH = pos ? p.rig : 0.0;
    r.Nb = (pos && bneg) ? p.rwg : 0.0;
    return r;
}

// d basal_drag / dN̄ for regularized Coulomb (τc = α² N̄)
__device__ __forceinline__ double drag_dN(const PhysK &p, double U, double C, double Nf) {
    if (p.friction != 2) return 0.0;
    const double m = p.m, tc = p.alpha2 * Nf;
    const double D = pow(C, 1.0 / m) * U + pow(tc, 1.0 / m);
    return p.alpha2 * C * pow(U, m - 1.0) * pow(D, -m - 1.0) * pow(C, 1.0 / m) * U;
}

// Pass 1, one thread per face: the H̄ and b̄ a u face sends to its cells (lo = the cell read as i − 1, hi = as i) and
// likewise for v faces (lo = j − 1, hi = j). β̄ is Az xs q̄ on active rows (SSA; xs the vector the operator multiplies,
// xu, xv the point the drag is linearized at), or the given basal β̄ᵇ (DIVA, with xu, xv the basal velocities).
__global__ __launch_bounds__(NT) void k_adj_geometry_faces(KG g, PhysK p, const uint8_t *__restrict__ act, const double *__restrict__ H,
                                                           const double *__restrict__ bed, const double *__restrict__ s,
                                                           const double *__restrict__ mu, const double *__restrict__ N,
                                                           const double *__restrict__ C, const double *__restrict__ phiu,
                                                           const double *__restrict__ phiv, const double *__restrict__ xu,
                                                           const double *__restrict__ xv, const double *__restrict__ lu,
                                                           const double *__restrict__ lv, const double *__restrict__ qbu,
                                                           const double *__restrict__ qbv, const double *__restrict__ bbar_u,
                                                           const double *__restrict__ bbar_v, const double *__restrict__ xsu,
                                                           const double *__restrict__ xsv, int gl, double *__restrict__ hlo_u,
                                                           double *__restrict__ hhi_u, double *__restrict__ blo_u,
                                                           double *__restrict__ bhi_u, double *__restrict__ hlo_v,
                                                           double *__restrict__ hhi_v, double *__restrict__ blo_v,
                                                           double *__restrict__ bhi_v) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy, ri = p.ri, rw = p.rw, gr = p.gr;
    for (int dir = 0; dir < 2; dir++) {
        const bool in = dir == 0 ? (i <= nxf(g) && j <= g.Ny) : (i <= g.Nx && j <= nyf(g));
        if (!in) continue;
        // the two cells this face reads, (a0, c0) below and (a1, c1) at the face's own index
        const int a0 = dir == 0 ? cx(g, i - 1) : cx(g, i), c0 = dir == 0 ? cy(g, j) : cy(g, j - 1);
        const int a1 = cx(g, i), c1 = cy(g, j);
        const long o0 = off(g, a0, c0), o1 = off(g, a1, c1);
        const double H0 = H[o0], H1 = H[o1], s0 = s[o0], s1 = s[o1], b0 = bed[o0], b1 = bed[o1];
        const double len = dir == 0 ? g.dy : g.dx, h = dir == 0 ? g.dx : g.dy;
        double hb0 = 0.0, hb1 = 0.0, bb0 = 0.0, bb1 = 0.0, sb0 = 0.0, sb1 = 0.0, nb0 = 0.0, nb1 = 0.0;
        const bool wall = dir == 0 ? periph_fc(g, i, j) : periph_cf(g, i, j);
        // driving stress, τ̄ = λ
        const double tb = dir == 0 ? lu[o] : lv[o];
        bool icefree = false, west = false, east = false;
        if (mu) {
            const double m0 = mu[o0], m1 = mu[o1];
            icefree = (m0 + m1) == 0.0;
            west = (m0 > 0) & !(m1 > 0);
            east = !(m0 > 0) & (m1 > 0);
        }
        if (!wall && !icefree && tb != 0.0) {
            bool st = false;
            if (gl) {
                const double f1 = H0 + rw / ri * b0, f2 = H1 + rw / ri * b1;
                st = ((f1 > 0) != (f2 > 0)) & (H0 > 0) & (H1 > 0);
            }
            if (west || east) {
                double dH, ds;
                if (west) { front_force_d(H0, s0, ri, rw, gr, dH, ds); hb0 += tb * len * dH; sb0 += tb * len * ds; }
                if (east) { front_force_d(H1, s1, ri, rw, gr, dH, ds); hb1 -= tb * len * dH; sb1 -= tb * len * ds; }
            } else if (st) {
                double d[6];
                gl_stress_d(H0, H1, s0, s1, b0, b1, ri, rw, gr, d);
                hb0 += tb * len * d[0]; hb1 += tb * len * d[1];
                sb0 += tb * len * d[2]; sb1 += tb * len * d[3];
                bb0 += tb * len * d[4]; bb1 += tb * len * d[5];
            } else {
                const double c = -Az * ri * gr, Hf = (H0 + H1) / 2.0, ds = (s1 - s0) / h;
                hb0 += tb * c * ds / 2.0;
                hb1 += tb * c * ds / 2.0;
                sb1 += tb * c * Hf / h;
                sb0 -= tb * c * Hf / h;
            }
        }
        // drag: β = φ · basal_drag(U, C̄, N̄)
        if (p.friction != 0 && !wall) {
            double bb;
            if (dir == 0) bb = bbar_u ? bbar_u[o] : qbar_u(g, act, qbu, i, j) * Az * xsu[o];
            else bb = bbar_v ? bbar_v[o] : qbar_v(g, act, qbv, i, j) * Az * xsv[o];
            if (bb != 0.0) {
                double U;
                if (dir == 0) {
                    const double vm = ((ldcf(g, xv, i - 1, j) + ldcf(g, xv, i, j)) / 2.0 + (ldcf(g, xv, i - 1, j + 1) + ldcf(g, xv, i, j + 1)) / 2.0) / 2.0;
                    U = sqrt(xu[o] * xu[o] + vm * vm + p.u0 * p.u0);
                } else {
                    const double um = ((ldfc(g, xu, i, j - 1) + ldfc(g, xu, i + 1, j - 1)) / 2.0 + (ldfc(g, xu, i, j) + ldfc(g, xu, i + 1, j)) / 2.0) / 2.0;
                    U = sqrt(xv[o] * xv[o] + um * um + p.u0 * p.u0);
                }
                const double Cf = (C[o0] + C[o1]) / 2.0;
                const double Nf = p.friction == 2 ? (N[o0] + N[o1]) / 2.0 : 0.0;
                const double ph = dir == 0 ? phiu[o] : phiv[o];
                if (p.grounding != 0) {
                    const double phb = bb * basal_drag(p, U, Cf, Nf);
                    double d0, d1;
                    segment_fraction_d(H0 + p.rw_ri * b0, H1 + p.rw_ri * b1, d0, d1);
                    hb0 += phb * d0; bb0 += phb * d0 * p.rw_ri;
                    hb1 += phb * d1; bb1 += phb * d1 * p.rw_ri;
                }
                if (p.friction == 2) {
                    const double nb = bb * ph * drag_dN(p, U, Cf, Nf) / 2.0;
                    nb0 += nb;
                    nb1 += nb;
                }
            }
        }
        // s̄ and N̄ of the two cells to H̄ and b̄ through their own geometry
        const GeoD g0 = geometry_d(p, H0, b0), g1 = geometry_d(p, H1, b1);
        hb0 += sb0 * g0.sH + nb0 * g0.NH;
        bb0 += sb0 * g0.sb + nb0 * g0.Nb;
        hb1 += sb1 * g1.sH + nb1 * g1.NH;
        bb1 += sb1 * g1.sb + nb1 * g1.Nb;
        if (dir == 0) { hlo_u[o] = hb0; hhi_u[o] = hb1; blo_u[o] = bb0; bhi_u[o] = bb1; }
        else { hlo_v[o] = hb0; hhi_v[o] = hb1; blo_v[o] = bb0; bhi_v[o] = bb1; }
    }
}

// Pass 2, one thread per cell (adding to H̄, b̄ when accumulate, e.g. over the steps of an unrolled Picard loop):
// H̄ = η̄H̄ μ Σ ησ Δζ (+ F̄₂ Σ w₂/ησ, DIVA) + the mask's share (∂μ/∂H · η̄H̄ H Σ ησ Δζ)
// + the sends of the faces that read the cell: along x the face at the cell's own index (read as i) and the one
// above it (read as i − 1), and on a Bounded edge the boundary face that the clamp maps here, on a Periodic one
// the first face, which wraps to the last cell; likewise along y. b̄ from the faces alone.
__global__ __launch_bounds__(NT) void k_adj_geometry_cells(KG g, PhysK p, Sig sg, const double *__restrict__ H,
                                                           const double *__restrict__ es, long plane, const double *__restrict__ etab,
                                                           const double *__restrict__ F2b, const double *__restrict__ hlo_u,
                                                           const double *__restrict__ hhi_u, const double *__restrict__ blo_u,
                                                           const double *__restrict__ bhi_u, const double *__restrict__ hlo_v,
                                                           const double *__restrict__ hhi_v, const double *__restrict__ blo_v,
                                                           const double *__restrict__ bhi_v, double *__restrict__ Hb,
                                                           double *__restrict__ Bb, int accumulate) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    double e = 0.0, w = 0.0;
    for (int k = 0; k < p.Ns; k++) {
        const double ek = es[o + (long)k * plane];
        e += ek * sg.dz[k];
        w += (ek == 0.0 ? 0.0 : sg.w2[k] / ek);
    }
    const double eb = etab[o];
    double mu = 1.0, dmu = 0.0;
    if (p.masking) {
        const double t = jclamp((H[o] - p.hmin) / p.htr, 0.0, 1.0);
        mu = t * t * (3.0 - 2.0 * t);
        dmu = (t > 0.0 && t < 1.0) ? 6.0 * t * (1.0 - t) / p.htr : 0.0;
    }
    double hb = eb * e * mu + eb * e * H[o] * dmu;
    if (F2b) hb += F2b[o] * w;
    double bb = 0.0;
    // x: face i (read as i), face i + 1 (read as i − 1), and the edge face mapped here
    hb += hhi_u[o]; bb += bhi_u[o];
    if (i + 1 <= nxf(g)) { const long q = off(g, i + 1, j); hb += hlo_u[q]; bb += blo_u[q]; }
    if (g.tx && i == g.Nx) { const long q = off(g, g.Nx + 1, j); hb += hhi_u[q]; bb += bhi_u[q]; }
    if ((g.tx && i == 1) || (!g.tx && i == g.Nx)) { const long q = off(g, 1, j); hb += hlo_u[q]; bb += blo_u[q]; }
    // y likewise
    hb += hhi_v[o]; bb += bhi_v[o];
    if (j + 1 <= nyf(g)) { const long q = off(g, i, j + 1); hb += hlo_v[q]; bb += blo_v[q]; }
    if (g.ty && j == g.Ny) { const long q = off(g, i, g.Ny + 1); hb += hhi_v[q]; bb += bhi_v[q]; }
    if ((g.ty && j == 1) || (!g.ty && j == g.Ny)) { const long q = off(g, i, 1); hb += hlo_v[q]; bb += blo_v[q]; }
    Hb[o] = accumulate ? Hb[o] + hb : hb;
    Bb[o] = accumulate ? Bb[o] + bb : bb;
}

}  // namespace nk
