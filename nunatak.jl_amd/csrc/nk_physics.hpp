// nk_physics.hpp — FP64 pointwise physics of the velocity solve: geometry and masks, driving stress with
// calving-front and grounding-line terms, Glen viscosity with its column integrals, basal drag, the DIVA
// coefficient chain and the post-solve velocity reconstruction. One thread per horizontal point; σ loops run
// inside the thread so each column is read once.
#pragma once
#include "nk_device.hpp"

namespace nk {

struct PhysK {
    double ri, rw, gr;            // ρᵢ, ρw, g
    double ri_rw, rw_ri, rig, rwg; // host-computed ratios and products, as the reference computes them
    int grounding, masking;
    double hmin, htr;
    double n, e0, inv_n, exp_e;   // −1/n and (1 − n)/2n computed on the host in FP64
    int A3d, friction;
    double m, alpha2, u0;
    int walls, Ns;
    double wf;                    // wall shear factor
};

// σ tables: zf[Ns+1], zc[Ns], dz[Ns], w2[Ns], pw[Ns*Ns] (profile weights, p = 1)
struct Sig {
    const double *zf, *zc, *dz, *w2, *pw;
};

// ---------------------------------------------------------------------------------------------------------------
// Geometry (src/Groundings/subgrid_grounded_fraction.jl:31-232, src/Models/smooth_ice_mask.jl:50-69). Bit-exact.
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double nonzero(double x) { return x == 0.0 ? 1.0 : x; }
__device__ __forceinline__ double ice_mask_value(double H, double hmin, double htr) {
    const double t = jclamp((H - hmin) / htr, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}
__device__ __forceinline__ double segment_fraction(double f1, double f2) {
    const double fp = jmax(f1, 0.0) + jmax(f2, 0.0);
    const double sf = fabs(f1) + fabs(f2);
    return sf == 0.0 ? 1.0 : fp / nonzero(sf);
}
__device__ __forceinline__ double triangle_fraction(double f1, double f2, double f3) {
    const double a = jmin(jmin(f1, f2), f3);
    const double c = jmax(jmax(f1, f2), f3);
    const double b = f1 + f2 + f3 - a - c;
    const double one_grounded = c * c / nonzero((c - a) * (c - b));
    const double one_floating = 1.0 - a * a / nonzero((b - a) * (c - a));
    const double phi = b > 0 ? one_floating : one_grounded;
    return a >= 0 ? 1.0 : (c <= 0 ? 0.0 : phi);
}

__global__ __launch_bounds__(NT) void k_geometry(KG g, PhysK p, const double *__restrict__ H, const double *__restrict__ b,
                                                 double *__restrict__ s, double *__restrict__ phi, double *__restrict__ phiu,
                                                 double *__restrict__ phiv, double *__restrict__ N, double *__restrict__ mu) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    const bool sub = p.grounding != 0;
    auto fl = [&](int a, int c) { return ldc(g, H, a, c) + p.rw_ri * ldc(g, b, a, c); };
    if (i <= g.Nx && j <= g.Ny) {
        const double Hc = H[o], bc = b[o];
        s[o] = sub ? jmax(bc + Hc, (1.0 - p.ri_rw) * Hc) : bc + Hc;
        if (sub) {
            const double fc = fl(i, j);
            const double fe = (fc + fl(i + 1, j)) / 2.0, fw = (fc + fl(i - 1, j)) / 2.0;
            const double fn = (fc + fl(i, j + 1)) / 2.0, fs = (fc + fl(i, j - 1)) / 2.0;
            const double fne = (fc + fl(i + 1, j) + fl(i, j + 1) + fl(i + 1, j + 1)) / 4.0;
            const double fnw = (fc + fl(i - 1, j) + fl(i, j + 1) + fl(i - 1, j + 1)) / 4.0;
            const double fse = (fc + fl(i + 1, j) + fl(i, j - 1) + fl(i + 1, j - 1)) / 4.0;
            const double fsw = (fc + fl(i - 1, j) + fl(i, j - 1) + fl(i - 1, j - 1)) / 4.0;
            const double t = triangle_fraction(fc, fe, fne) + triangle_fraction(fc, fn, fne) + triangle_fraction(fc, fn, fnw) +
                             triangle_fraction(fc, fw, fnw) + triangle_fraction(fc, fw, fsw) + triangle_fraction(fc, fs, fsw) +
                             triangle_fraction(fc, fs, fse) + triangle_fraction(fc, fe, fse);
            phi[o] = t / 8.0;
        } else {
            phi[o] = 1.0;
        }
        const double P0 = p.rig * Hc;
        N[o] = jmax(P0 + p.rwg * jmin(bc, 0.0), 0.0);
        mu[o] = ice_mask_value(Hc, p.hmin, p.htr);
    }
    if (i <= nxf(g) && j <= g.Ny) phiu[o] = sub ? segment_fraction(fl(i - 1, j), fl(i, j)) : 1.0;
    if (i <= g.Nx && j <= nyf(g)) phiv[o] = sub ? segment_fraction(fl(i, j - 1), fl(i, j)) : 1.0;
}

// ---------------------------------------------------------------------------------------------------------------
// Driving stress (src/StressBalances/driving_stress.jl:56-125). Bit-exact, with the face classes.
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double front_force(double Hi, double si, double ri, double rw, double gr) {
    const double D = jclamp(Hi - si, 0.0, Hi);
    return gr * (ri * (Hi * Hi) - rw * (D * D)) / 2.0;
}
__device__ __forceinline__ double gl_stress(double H1, double H2, double s1, double s2, double b1, double b2, double ri, double rw,
                                            double gr, bool &st) {
    const double f1 = H1 + rw / ri * b1;
    const double f2 = H2 + rw / ri * b2;
    st = ((f1 > 0) != (f2 > 0)) & (H1 > 0) & (H2 > 0);
    const double theta = f1 / (st ? f1 - f2 : 1.0);
    const double Hg = H1 + theta * (H2 - H1);
    const double sg = (1.0 - ri / rw) * Hg;
    return -ri * gr * ((H1 + Hg) * (sg - s1) + (Hg + H2) * (s2 - sg)) / 2.0;
}

__global__ __launch_bounds__(NT) void k_driving_stress(KG g, PhysK p, const double *__restrict__ H, const double *__restrict__ s,
                                                       const double *__restrict__ mu, const double *__restrict__ b,
                                                       double *__restrict__ tu, double *__restrict__ tv, uint8_t *__restrict__ cu,
                                                       uint8_t *__restrict__ cv) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    const double Az = g.dx * g.dy;
    const double ri = p.ri, rw = p.rw, gr = p.gr;
    if (i <= nxf(g) && j <= g.Ny) {
        const double H0 = ldc(g, H, i - 1, j), H1 = ldc(g, H, i, j), s0 = ldc(g, s, i - 1, j), s1 = ldc(g, s, i, j);
        bool icefree = false, front = false, st = false;
        double tf = 0.0, tg = 0.0;
        if (mu) {
            const double m0 = ldc(g, mu, i - 1, j), m1 = ldc(g, mu, i, j);
            icefree = (m0 + m1) == 0.0;
            const bool west = (m0 > 0) & !(m1 > 0), east = !(m0 > 0) & (m1 > 0);
            const double Fv = (west ? front_force(H0, s0, ri, rw, gr) : 0.0) - (east ? front_force(H1, s1, ri, rw, gr) : 0.0);
            front = west | east;
            tf = g.dy * Fv;
        }
        if (b) tg = g.dy * gl_stress(H0, H1, s0, s1, ldc(g, b, i - 1, j), ldc(g, b, i, j), ri, rw, gr, st);
        const double Hf = (H0 + H1) / 2.0;
        const double ds = (s1 - s0) / g.dx;
        const double t = -Az * ri * gr * Hf * ds;
        const bool wall = periph_fc(g, i, j);
        tu[o] = (wall || icefree) ? 0.0 : (front ? tf : (st ? tg : t));
        cu[o] = (uint8_t)((wall ? 1 : 0) | (icefree ? 2 : 0) | (front ? 4 : 0) | (st ? 8 : 0));
    }
    if (i <= g.Nx && j <= nyf(g)) {
        const double H0 = ldc(g, H, i, j - 1), H1 = ldc(g, H, i, j), s0 = ldc(g, s, i, j - 1), s1 = ldc(g, s, i, j);
        bool icefree = false, front = false, st = false;
        double tf = 0.0, tg = 0.0;
        if (mu) {
            const double m0 = ldc(g, mu, i, j - 1), m1 = ldc(g, mu, i, j);
            icefree = (m0 + m1) == 0.0;
            const bool south = (m0 > 0) & !(m1 > 0), north = !(m0 > 0) & (m1 > 0);
            const double Fv = (south ? front_force(H0, s0, ri, rw, gr) : 0.0) - (north ? front_force(H1, s1, ri, rw, gr) : 0.0);
            front = south | north;
            tf = g.dx * Fv;
        }
        if (b) tg = g.dx * gl_stress(H0, H1, s0, s1, ldc(g, b, i, j - 1), ldc(g, b, i, j), ri, rw, gr, st);
        const double Hf = (H0 + H1) / 2.0;
        const double ds = (s1 - s0) / g.dy;
        const double t = -Az * ri * gr * Hf * ds;
        const bool wall = periph_cf(g, i, j);
        tv[o] = (wall || icefree) ? 0.0 : (front ? tf : (st ? tg : t));
        cv[o] = (uint8_t)((wall ? 1 : 0) | (icefree ? 2 : 0) | (front ? 4 : 0) | (st ? 8 : 0));
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Viscosity and column integrals (src/Rheologies/glen_flow_law.jl:97-143, src/StressBalances/column_integrals.jl:54-72)
// ---------------------------------------------------------------------------------------------------------------

// effective_strain_rate² (src/StressBalances/strain_rates.jl:27-32): strain rates see walls only.
__device__ __forceinline__ double eff_strain2(const KG &g, const double *u, const double *v, int i, int j, double wf) {
    const double exx = (iu_w(g, u, i + 1, j) - iu_w(g, u, i, j)) / g.dx;
    const double eyy = (iv_w(g, v, i, j + 1) - iv_w(g, v, i, j)) / g.dy;
    auto exy = [&](int a, int c) {
        double e = ((iu_w(g, u, a, c) - iu_w(g, u, a, c - 1)) / g.dy + (iv_w(g, v, a, c) - iv_w(g, v, a - 1, c)) / g.dx) / 2.0;
        return periph_ff(g, a, c) ? wf * e : e;
    };
    const double A = exy(i, j), B = exy(i + 1, j), C = exy(i, j + 1), D = exy(i + 1, j + 1);
    const double exy2 = ((A * A + B * B) / 2.0 + (C * C + D * D) / 2.0) / 2.0;
    return exx * exx + eyy * eyy + exx * eyy + exy2;
}

// mode 0: SSA viscosity; mode 1: DIVA viscosity with the lagged vertical shear. Both fuse the column integrals.
__global__ __launch_bounds__(NT) void k_viscosity(KG g, PhysK p, Sig sg, int mode, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ A, long plane,
                                                  double *__restrict__ es, const double *__restrict__ H, const double *__restrict__ mu,
                                                  const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                  const double *__restrict__ ub, const double *__restrict__ vb,
                                                  double *__restrict__ eta, double *__restrict__ F2) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double e2m = eff_strain2(g, u, v, i, j, p.wf);
    double tx = 0.0, ty = 0.0;
    if (mode == 1) {
        tx = (ldfc(g, bbu, i, j) * ldfc(g, ub, i, j) + ldfc(g, bbu, i + 1, j) * ldfc(g, ub, i + 1, j)) / 2.0;
        ty = (ldcf(g, bbv, i, j) * ldcf(g, vb, i, j) + ldcf(g, bbv, i, j + 1) * ldcf(g, vb, i, j + 1)) / 2.0;
    }
    double pa2 = 0.0, pe = 0.0;
    if (!p.A3d) pa2 = pow(A[o], p.inv_n);
    if (mode == 0) pe = pow(e2m + p.e0 * p.e0, p.exp_e);
    double et = 0.0, Fs = 0.0;
    for (int k = 0; k < p.Ns; k++) {
        const long ok = o + (long)k * plane;
        const double pa = p.A3d ? pow(A[ok], p.inv_n) : pa2;
        double pk;
        if (mode == 0) {
            pk = pe;
        } else {
            const double zeta = sg.zc[k];
            double em = es[ok];
            em = em == 0.0 ? INFINITY : em;
            const double exz = tx * (1.0 - zeta) / (2.0 * em);
            const double eyz = ty * (1.0 - zeta) / (2.0 * em);
            pk = pow(e2m + exz * exz + eyz * eyz + p.e0 * p.e0, p.exp_e);
        }
        const double ek = pa * pk / 2.0;
        es[ok] = ek;
        et += ek * sg.dz[k];
        Fs += 1.0 / ek * sg.w2[k];
    }
    eta[o] = et * H[o] * (mu ? mu[o] : 1.0);
    F2[o] = Fs * H[o];
}

// basal_drag (src/Frictions/frictions.jl:83-89)
__device__ __forceinline__ double basal_drag(const PhysK &p, double U, double Cf, double Nf) {
    if (p.friction == 1) return Cf * pow(U, p.m - 1.0);
    const double tc = p.alpha2 * Nf;
    return Cf * pow(U, p.m - 1.0) * tc / pow(pow(Cf, 1.0 / p.m) * U + pow(tc, 1.0 / p.m), p.m);
}

// compute_basal_drag! (src/Frictions/frictions.jl:127-160): β from velocities (u, v) at faces, zero at walls.
__global__ __launch_bounds__(NT) void k_drag(KG g, PhysK p, const double *__restrict__ u, const double *__restrict__ v,
                                             const double *__restrict__ C, const double *__restrict__ N,
                                             const double *__restrict__ phiu, const double *__restrict__ phiv,
                                             double *__restrict__ bu, double *__restrict__ bv) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    const bool useN = p.friction == 2;
    const double u0 = p.u0;
    if (i <= nxf(g) && j <= g.Ny) {
        const double vm = ((ldcf(g, v, i - 1, j) + ldcf(g, v, i, j)) / 2.0 + (ldcf(g, v, i - 1, j + 1) + ldcf(g, v, i, j + 1)) / 2.0) / 2.0;
        const double uu = u[o];
        const double U = sqrt(uu * uu + vm * vm + u0 * u0);
        const double Cf = (ldc(g, C, i - 1, j) + ldc(g, C, i, j)) / 2.0;
        const double Nf = useN ? (ldc(g, N, i - 1, j) + ldc(g, N, i, j)) / 2.0 : 0.0;
        const double beta = phiu[o] * basal_drag(p, U, Cf, Nf);
        bu[o] = periph_fc(g, i, j) ? 0.0 : beta;
    }
    if (i <= g.Nx && j <= nyf(g)) {
        const double um = ((ldfc(g, u, i, j - 1) + ldfc(g, u, i + 1, j - 1)) / 2.0 + (ldfc(g, u, i, j) + ldfc(g, u, i + 1, j)) / 2.0) / 2.0;
        const double vv = v[o];
        const double U = sqrt(vv * vv + um * um + u0 * u0);
        const double Cf = (ldc(g, C, i, j - 1) + ldc(g, C, i, j)) / 2.0;
        const double Nf = useN ? (ldc(g, N, i, j - 1) + ldc(g, N, i, j)) / 2.0 : 0.0;
        const double beta = phiv[o] * basal_drag(p, U, Cf, Nf);
        bv[o] = periph_cf(g, i, j) ? 0.0 : beta;
    }
}

// DIVA: uᵇ = u / (1 + βᵇ ℑx(F₂)) (src/StressBalances/diva.jl:1-27) over the face extents
__global__ __launch_bounds__(NT) void k_basal_velocity(KG g, const double *__restrict__ u, const double *__restrict__ v,
                                                       const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                       const double *__restrict__ F2, double *__restrict__ ub, double *__restrict__ vb) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    if (i <= nxf(g) && j <= g.Ny) ub[o] = u[o] / (1.0 + bbu[o] * ((ldc(g, F2, i - 1, j) + ldc(g, F2, i, j)) / 2.0));
    if (i <= g.Nx && j <= nyf(g)) vb[o] = v[o] / (1.0 + bbv[o] * ((ldc(g, F2, i, j - 1) + ldc(g, F2, i, j)) / 2.0));
}

// DIVA: β = βᵇ / (1 + βᵇ ℑx(F₂)) (diva.jl:35-51)
__global__ __launch_bounds__(NT) void k_effective_drag(KG g, const double *__restrict__ bbu, const double *__restrict__ bbv,
                                                       const double *__restrict__ F2, double *__restrict__ bu, double *__restrict__ bv) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    if (i <= nxf(g) && j <= g.Ny) bu[o] = bbu[o] / (1.0 + bbu[o] * ((ldc(g, F2, i - 1, j) + ldc(g, F2, i, j)) / 2.0));
    if (i <= g.Nx && j <= nyf(g)) bv[o] = bbv[o] / (1.0 + bbv[o] * ((ldc(g, F2, i, j - 1) + ldc(g, F2, i, j)) / 2.0));
}

// compute_velocity_profile! (diva.jl:89-109): uσ = uᵇ (1 + βᵇ ℑx(F₁)), F₁ = H ∫₀^ζₖ (1 − ζ′)/η dζ′
__global__ __launch_bounds__(NT) void k_velocity_profile(KG g, Sig sg, int Ns, long plane, const double *__restrict__ es,
                                                         const double *__restrict__ H, const double *__restrict__ ub,
                                                         const double *__restrict__ vb, const double *__restrict__ bbu,
                                                         const double *__restrict__ bbv, double *__restrict__ us,
                                                         double *__restrict__ vs) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    auto shear_integral = [&](int a, int c, int k) {
        const long oc = off(g, cx(g, a), cy(g, c));
        double I = 0.0;
        for (int kp = 0; kp <= k; kp++) I += 1.0 / es[oc + (long)kp * plane] * sg.pw[k * Ns + kp];
        return I * H[oc];
    };
    for (int k = 0; k < Ns; k++) {
        if (i <= nxf(g) && j <= g.Ny) {
            const double F1 = (shear_integral(i - 1, j, k) + shear_integral(i, j, k)) / 2.0;
            us[o + (long)k * plane] = ub[o] * (1.0 + bbu[o] * F1);
        }
        if (i <= g.Nx && j <= nyf(g)) {
            const double F1 = (shear_integral(i, j - 1, k) + shear_integral(i, j, k)) / 2.0;
            vs[o + (long)k * plane] = vb[o] * (1.0 + bbv[o] * F1);
        }
    }
}

// extend_depth_average! (src/Models/ice_sheet_model_methods.jl:167-180)
__global__ __launch_bounds__(NT) void k_extend(KG g, int Ns, long plane, const double *__restrict__ u, const double *__restrict__ v,
                                               double *__restrict__ ub, double *__restrict__ vb, double *__restrict__ us,
                                               double *__restrict__ vs) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    ub[o] = u[o];
    vb[o] = v[o];
    for (int k = 0; k < Ns; k++) {
        us[o + (long)k * plane] = u[o];
        vs[o + (long)k * plane] = v[o];
    }
}

// Arrhenius rate factor from temperature through the enthalpy round trip
// (src/Thermodynamics/enthalpy_conversions.jl:1-66, src/Thermodynamics/column_energy_solve.jl:105-117,
//  src/Rheologies/arrhenius_rate_factor.jl:102-110), default parameters of the reference.
// Defaults of EnthalpyThermodynamics (src/Thermodynamics/enthalpy_thermodynamics.jl:96-105) and ArrheniusRateFactor
// (src/Rheologies/arrhenius_rate_factor.jl:60-69).
struct Thermo {
    static constexpr double ci = 2009.0, L = 3.34e5, be = 7.9e-8, T0 = 273.15, Tr = 223.15;
    static constexpr double Ac = 3.61e-13, Aw = 1.73e3, Qc = 6.0e4, Qw = 13.9e4, Tt = 263.15, Tw = 1.0, R = 8.31441,
                            cw = 181.25, wmax = 0.01;
};
// temperature / water_fraction of enthalpy E at pressure p (src/Thermodynamics/enthalpy_conversions.jl:35-52), then
// arrhenius_rate_factor(T − Tₚₘ + T₀, ω) (arrhenius_rate_factor.jl:102-110, column_energy_solve.jl:106-111)
__device__ __forceinline__ void thermal_state(double E, double pr, double &Tk, double &w, double &A) {
    using C = Thermo;
    Tk = jmin(C::T0 - C::be * pr, C::Tr + E / C::ci);
    const double Epm = C::ci * ((C::T0 - C::be * pr) - C::Tr);
    w = jmax(0.0, E - Epm) / C::L;
    const double Tp = Tk - (C::T0 - C::be * pr) + C::T0;
    const double lc = log(C::Ac) - C::Qc / (C::R * Tp);
    const double lw = log(C::Aw) - C::Qw / (C::R * Tp);
    const double x = jclamp((Tp - C::Tt) / C::Tw + 1.0 / 2.0, 0.0, 1.0);
    const double s = x * x * (3.0 - 2.0 * x);
    const double soft = 1.0 + C::cw * jmin(w, C::wmax);
    A = exp((1.0 - s) * lc + s * lw) * soft;
}

// A from TEMPERATURE through the enthalpy round trip of set!(model; T) (ice_sheet_model_methods.jl:47-57)
__global__ __launch_bounds__(NT) void k_rate_factor(KG g, Sig sg, int Ns, long plane, double rig, const double *__restrict__ T,
                                                    const double *__restrict__ H, double *__restrict__ A, double *__restrict__ Eout) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    using C = Thermo;
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * plane;
        const double pr = rig * H[o] * (1.0 - sg.zc[k]);
        const double Tpm = C::T0 - C::be * pr;
        const double Tc = jmin(T[ok], Tpm);
        const double E = C::ci * (Tc - C::Tr) + 0.0 * C::L;
        Eout[ok] = E;
        double Tk, w;
        thermal_state(E, pr, Tk, w, A[ok]);
    }
}

// update_thermal_diagnostics! (column_energy_solve.jl:85-117): T, ω and A from the enthalpy E of every layer
__global__ __launch_bounds__(NT) void k_thermal_diagnostics(KG g, Sig sg, int Ns, long plane, double rig, const double *__restrict__ E,
                                                            const double *__restrict__ H, double *__restrict__ T,
                                                            double *__restrict__ W, double *__restrict__ A) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    for (int k = 0; k < Ns; k++) {
        const long ok = o + (long)k * plane;
        const double pr = rig * H[o] * (1.0 - sg.zc[k]);
        thermal_state(E[ok], pr, T[ok], W[ok], A[ok]);
    }
}

}  // namespace nk
