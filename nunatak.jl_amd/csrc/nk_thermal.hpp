// nk_thermal.hpp — the enthalpy step on the device (src/Thermodynamics/column_energy_solve.jl, layer_transport.jl,
// strain_heating.jl): strain heating, then per column the layer volume fluxes, the basal state and the implicit
// tridiagonal system of the layers (k_energy_assemble, which reads the neighbours' old enthalpy for the explicit
// upwind transport), then the column solve, drainage, basal water and the thermal diagnostics (k_energy_solve).
// One thread per column; the σ loops run inside the thread.
#pragma once
#include "nk_device.hpp"
#include "nk_physics.hpp"
#include "../../include/nunatak_hip.h"

namespace nk {

// EnthalpyParameters defaults (src/Thermodynamics/enthalpy_thermodynamics.jl:151-166)
struct EP {
    static constexpr double ci = 2009.0, kc = 2.1, L = 3.34e5, beta = 7.9e-8, T0 = 273.15, Tr = 223.15;
    static constexpr double ratio = 1e-3, dE = 1.0, hwb = 1e-3, Hmin = 1.0, wd = 0.01, rmax = 1.58443823077064e-9;
};

struct EnergyArgs {
    double rho, grav, dt;
    int grounding, coupled, Ns, a3d, till;
    double till_rate, till_wmax;   // drain_till (till_effective_pressure.jl:77-78) when till
    long plane;
    const double *H, *Hn, *us, *vs, *ub, *vb, *bu, *bv, *phi, *m, *G, *Ts, *To, *Phi;
    double *E, *Hw, *mg, *T, *W, *A, *W3, *a, *b, *c, *f, *t, *bas;   // W3: Ns + 1 face planes; bas: rᵇ, sᵇ, Dᵇ, qᵍ
};

__device__ __forceinline__ double colthick(double H) { return jmax(H, EP::Hmin); }
__device__ __forceinline__ double unit_ramp(double x, double w) { return jclamp(x / w, 0.0, 1.0); }
__device__ __forceinline__ double pmT(double p) { return EP::T0 - EP::beta * p; }
__device__ __forceinline__ double pmE(double p) { return EP::ci * (pmT(p) - EP::Tr); }
__device__ __forceinline__ double enth(double T, double w, double p) { return EP::ci * (jmin(T, pmT(p)) - EP::Tr) + w * EP::L; }
// layer_diffusivity (src/Thermodynamics/enthalpy_conversions.jl:72-78)
__device__ __forceinline__ double layer_diff(double E, double p) {
    const double w = unit_ramp(E - pmE(p) + EP::dE, EP::dE);
    const double Kc = EP::kc / EP::ci;
    return Kc * (1.0 - w + w * EP::ratio);
}
__device__ __forceinline__ double upwind(double u, double cm, double cp) { return u > 0.0 ? cm : cp; }

// Strain heating (strain_heating.jl): Φ = 4 ησ ε̇ₑ² (SSA); DIVA adds τᵇ² w₂ₖ / Δζₖ / ησ.
__global__ __launch_bounds__(NT) void k_strain_heating(KG g, Sig sg, int Ns, long plane, double wf, int diva, const double *__restrict__ u,
                                                       const double *__restrict__ v, const double *__restrict__ es,
                                                       const double *__restrict__ ub, const double *__restrict__ vb,
                                                       const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                       double *__restrict__ Phi) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double e2 = eff_strain2(g, u, v, i, j, wf);
    double tb2 = 0.0;
    if (diva) {
        auto s2 = [](double b, double x) { const double t = b * x; return t * t; };
        tb2 = (s2(ldfc(g, bbu, i, j), ldfc(g, ub, i, j)) + s2(ldfc(g, bbu, i + 1, j), ldfc(g, ub, i + 1, j))) / 2.0 +
              (s2(ldcf(g, bbv, i, j), ldcf(g, vb, i, j)) + s2(ldcf(g, bbv, i, j + 1), ldcf(g, vb, i, j + 1))) / 2.0;
    }
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * plane;
        const double eta = es[ok];
        double val = 4.0 * eta * e2;
        if (diva) val = val + tb2 * sg.w2[k] / sg.dz[k] / (eta == 0.0 ? INFINITY : eta);
        Phi[ok] = val;
    }
}

// Per column: layer volume fluxes W (layer_transport.jl:34-44), basal state (column_energy_solve.jl:130-147) and the
// tridiagonal system of the layers (:185-226). Reads the old enthalpy of the neighbours; writes W3, bas, a, b, c, f.
__global__ __launch_bounds__(NT) void k_energy_assemble(KG g, Sig sg, EnergyArgs e) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = e.plane;
    const int Ns = e.Ns;
    const double rho = e.rho, Az = g.dx * g.dy;
    const double Hc = e.H[o], Ht = colthick(Hc), Htn = colthick(e.Hn[o]);
    const bool mu = Hc > EP::Hmin;
    const double phg = e.grounding ? e.phi[o] : 1.0;
    // layer volume fluxes through the faces of each layer
    const double dHdt = (Ht - Htn) / e.dt;
    const double bm = e.grounding ? (1.0 - e.phi[o]) * e.m[o] : e.m[o];
    double Wk = -bm - (e.coupled ? e.mg[o] : 0.0);
    e.W3[o] = Wk;
    for (int k = 0; k < Ns; k++) {
        const double *usk = e.us + (long)k * P, *vsk = e.vs + (long)k * P;
        auto xv = [&](int a) {
            const double u = periph_fc(g, a, j) ? 0.0 : ldfc(g, usk, a, j);
            return g.dy * u * upwind(u, ldc(g, e.Hn, a - 1, j), ldc(g, e.Hn, a, j));
        };
        auto yv = [&](int c) {
            const double v = periph_cf(g, i, c) ? 0.0 : ldcf(g, vsk, i, c);
            return g.dx * v * upwind(v, ldc(g, e.Hn, i, c - 1), ldc(g, e.Hn, i, c));
        };
        const double vfd = (xv(i + 1) - xv(i) + yv(j + 1) - yv(j)) / Az;
        Wk = Wk - sg.dz[k] * (dHdt + vfd);
        e.W3[o + (long)(k + 1) * P] = Wk;
    }
    // basal state
    const double pb = rho * e.grav * Hc;
    const double p1 = rho * e.grav * Hc * (1.0 - sg.zc[0]);
    const double E1 = e.E[o];
    const double D = layer_diff(E1, p1) / (Ht * sg.zc[0]);
    const double F = (ldfc(g, e.bu, i, j) * (ldfc(g, e.ub, i, j) * ldfc(g, e.ub, i, j)) +
                      ldfc(g, e.bu, i + 1, j) * (ldfc(g, e.ub, i + 1, j) * ldfc(g, e.ub, i + 1, j)) +
                      ldcf(g, e.bv, i, j) * (ldcf(g, e.vb, i, j) * ldcf(g, e.vb, i, j)) +
                      ldcf(g, e.bv, i, j + 1) * (ldcf(g, e.vb, i, j + 1) * ldcf(g, e.vb, i, j + 1))) / 2.0;
    const double Epmb = pmE(pb);
    const double Eb = E1 + (e.G[o] + F) / D;
    const double r = jmax(unit_ramp(Eb - Epmb, EP::dE), unit_ramp(e.Hw[o], EP::hwb));
    const double sb = unit_ramp(E1 - pmE(p1) + EP::dE, EP::dE);
    const double qg = phg * e.G[o] + F;
    e.bas[o] = r; e.bas[o + P] = sb; e.bas[o + 2 * P] = D; e.bas[o + 3 * P] = qg;
    // boundary terms
    const double Es = enth(e.Ts[o], 0.0, 0.0);
    const double chi = phg * r * (1.0 - sb);
    const double Eo = enth(e.To[o], 0.0, pb);
    const double Fb = rho * e.W3[o];
    const double Lam = D * (chi + 1.0 - phg) + jmax(-Fb, 0.0);
    const double S = (1.0 - r) * qg + D * (chi * Epmb + (1.0 - phg) * Eo) + jmax(Fb, 0.0) * Epmb;
    const double Fs = rho * e.W3[o + (long)Ns * P];
    const double pN = rho * e.grav * Hc * (1.0 - sg.zc[Ns - 1]);
    const double Ks = layer_diff(e.E[o + (long)(Ns - 1) * P], pN) / (Ht * (1.0 - sg.zc[Ns - 1]));
    const double Ds = Ks + jmax(Fs, 0.0);
    const double Ss = (Ks + jmax(-Fs, 0.0)) * Es;
    // couplings of interior face kf (between layers kf − 1 and kf, 0-based): (entry of E[kf−1] in row kf, of E[kf] in row kf−1)
    auto face = [&](int kf, double &lower, double &upper) {
        const double zf = sg.zf[kf], zm = sg.zc[kf - 1], zp = sg.zc[kf];
        const double th = (zf - zm) / (zp - zm);
        const double Fk = rho * e.W3[o + (long)kf * P];
        const double Km = layer_diff(e.E[o + (long)(kf - 1) * P], rho * e.grav * Hc * (1.0 - zm));
        const double Kp = layer_diff(e.E[o + (long)kf * P], rho * e.grav * Hc * (1.0 - zp));
        const double cond = 1.0 / (Ht * ((zf - zm) / Km + (zp - zf) / Kp));
        const double Dk = jmax(jmax(cond, jmax(Fk, 0.0) * th), jmax(-Fk, 0.0) * (1.0 - th));
        lower = -Fk * (1.0 - th) - Dk;
        upper = Fk * th - Dk;
    };
    double up_m = 0.0;   // coupling of E[k] in row k − 1 at face k (the previous iteration's upper⁺)
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * P;
        const double dzk = sg.dz[k];
        double lo_p = 0.0, up_p = 0.0;
        if (k + 1 < Ns) face(k + 1, lo_p, up_p);
        const double dm = k == 0 ? Lam : -up_m;
        const double dp = k == Ns - 1 ? Ds : -lo_p;
        const double forcing = (k == 0 ? S : 0.0) + (k == Ns - 1 ? Ss : 0.0);
        // enthalpy_flux_divergence (layer_transport.jl:25-28) with Hⁿ and the old enthalpy
        const double *usk = e.us + (long)k * P, *vsk = e.vs + (long)k * P, *Ek = e.E + (long)k * P;
        auto xe = [&](int a) {
            const double uraw = ldfc(g, usk, a, j);
            const double u = periph_fc(g, a, j) ? 0.0 : uraw;
            const double vol = g.dy * u * upwind(u, ldc(g, e.Hn, a - 1, j), ldc(g, e.Hn, a, j));
            return vol * upwind(uraw, ldc(g, Ek, a - 1, j), ldc(g, Ek, a, j));
        };
        auto ye = [&](int c) {
            const double vraw = ldcf(g, vsk, i, c);
            const double v = periph_cf(g, i, c) ? 0.0 : vraw;
            const double vol = g.dx * v * upwind(v, ldc(g, e.Hn, i, c - 1), ldc(g, e.Hn, i, c));
            return vol * upwind(vraw, ldc(g, Ek, i, c - 1), ldc(g, Ek, i, c));
        };
        const double efd = (xe(i + 1) - xe(i) + ye(j + 1) - ye(j)) / Az;
        const double transport = rho * dzk * efd;
        e.a[ok] = mu ? lo_p : 0.0;
        e.c[ok] = mu ? up_p : 0.0;
        e.b[ok] = mu ? rho * Ht * dzk / e.dt + dm + dp : 1.0;
        e.f[ok] = mu ? rho * Htn * dzk * e.E[ok] / e.dt - transport + Ht * dzk * e.Phi[ok] + forcing : Es;
        up_m = up_p;
    }
}

// N with a saturated-till law (till_effective_pressure.jl:66-73, subgrid_grounded_fraction.jl:176-183)
__global__ __launch_bounds__(NT) void k_till_pressure(KG g, double rig, double rwg, nk_till t, const double *__restrict__ H,
                                                      const double *__restrict__ b, const double *__restrict__ Hw,
                                                      double *__restrict__ N) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double P0 = rig * H[o];
    const double No = jmax(P0 + rwg * jmin(b[o], 0.0), 0.0);
    const double N0 = t.reference_effective_pressure, d = t.overburden_fraction;
    const double s = jclamp(Hw[o] / t.maximum_water_thickness, 0.0, 1.0);
    const double Pt = P0 > 0.0 ? P0 : N0 / d;
    const double Nt = N0 * pow(d * Pt / N0, s) * pow(10.0, t.reference_void_ratio / t.compressibility * (1.0 - s));
    N[o] = jmin(Nt, No);
}

// drainage_rate (column_energy_solve.jl:230-234)
__device__ __forceinline__ double drainage_rate(double w) {
    const double x = w / EP::wd;
    const double shape = jclamp(x - 1.0, 0.0, 1.0) / 10.0 + 9.0 * jclamp(x - 2.0, 0.0, 1.0) / 10.0;
    return EP::rmax * shape;
}

// Per column: the tridiagonal solve (Oceananigans BatchedTridiagonalSolver: Thomas algorithm, forward pass stopped
// where the pivot is not above 10 eps), drainage and basal water (_update_basal_water!, :238-267), the thermal
// diagnostics (:85-117) and Hⁿ = H.
__global__ __launch_bounds__(NT) void k_energy_solve(KG g, Sig sg, EnergyArgs e, double *__restrict__ Hn_out) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = e.plane;
    const int Ns = e.Ns;
    double beta = e.b[o];
    e.E[o] = e.f[o] / beta;
    for (int k = 1; k < Ns; k++) {
        const long ok = o + (long)k * P, om = ok - P;
        const double tk = e.c[om] / beta;
        e.t[ok] = tk;
        beta = e.b[ok] - e.a[om] * tk;
        if (!(fabs(beta) > 10.0 * 2.220446049250313e-16)) break;
        e.E[ok] = (e.f[ok] - e.a[om] * e.E[om]) / beta;
    }
    for (int k = Ns - 2; k >= 0; k--) {
        const long ok = o + (long)k * P;
        e.E[ok] = e.E[ok] - e.t[ok + P] * e.E[ok + P];
    }
    // melt, drainage and the basal water layer
    const double rho = e.rho, L = EP::L;
    const double Hc = e.H[o], Ht = colthick(Hc);
    const bool mu = Hc > EP::Hmin;
    const double phg = e.grounding ? e.phi[o] : 1.0;
    const double r = e.bas[o], sb = e.bas[o + P], D = e.bas[o + 2 * P], qg = e.bas[o + 3 * P];
    const double Epmb = pmE(rho * e.grav * Hc);
    const double q = D * (Epmb - e.E[o]);
    const double m = r * (qg - phg * (1.0 - sb) * q) / (rho * L);
    double drained = 0.0;
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * P;
        const double p = rho * e.grav * Hc * (1.0 - sg.zc[k]);
        const double w = jmax(0.0, e.E[ok] - pmE(p)) / L;
        double dw = jmin(e.dt * drainage_rate(w), jmax(w - EP::wd, 0.0));
        dw = mu ? dw : 0.0;
        e.E[ok] = e.E[ok] - L * dw;
        drained = drained + Ht * sg.dz[k] * dw;
    }
    const double Hw0 = e.Hw[o];
    const double Hstar = Hw0 + e.dt * (mu ? m : 0.0) + drained;
    const double Hplus = jmax(0.0, Hstar);
    const double missing = Hplus - Hstar;
    e.mg[o] = (Hplus - Hw0) / e.dt;
    e.Hw[o] = e.till ? jclamp(Hplus - e.dt * e.till_rate, 0.0, e.till_wmax) : Hplus;
    e.E[o] = e.E[o] - L * missing / (Ht * sg.dz[0]);
    // thermal diagnostics
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * P;
        const double p = rho * e.grav * Hc * (1.0 - sg.zc[k]);
        double Tk, w, A;
        thermal_state(e.E[ok], p, Tk, w, A);
        e.T[ok] = Tk;
        e.W[ok] = w;
        if (e.a3d) e.A[ok] = A;
    }
    Hn_out[o] = Hc;
}

}  // namespace nk
