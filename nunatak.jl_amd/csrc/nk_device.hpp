// nk_device.hpp — gfx950 device code of libnunatak_hip.so: stencil kernels of the velocity solve.
//
// Layout in HBM. Every horizontal field of a grid level is one pitched 2D array of (Ny + 1) rows of `pitch`
// elements (pitch = Nx + 1 rounded up to 64 elements = 512 B in FP64), element (i, j) (1-based, i fastest) at
// (j - 1) * pitch + (i - 1); σ fields stack Nσ such planes. There are no halo cells: neighbour indices are mapped
// in the kernels with the reference's halo semantics (Periodic wraps; Bounded Center clamps, which equals
// Oceananigans' zero-gradient halo fill; Face values beyond a Bounded wall are peripheral and never used unmasked).
//
// Arithmetic. Kernels are compiled with -ffp-contract=off and keep the reference's expression order, so the
// FP64 operator, masks, grounded fractions, driving stress and multigrid transfers reproduce the reference's
// IEEE results operation for operation. Reductions (dot products) are deterministic: each block writes a partial
// in a fixed slot and the last block to finish sums the slots in a fixed order, then runs a "finisher" that
// updates the PCG scalars on the device, so no host round trip is needed inside an iteration.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <stdint.h>

namespace nk {

constexpr int BX = 64;   // threads along x per block: one wavefront per row segment
constexpr int BY = 4;    // rows per block
constexpr int NT = BX * BY;
constexpr int MAXRED = 4;  // values reduced by one kernel at most

struct KG {
    int Nx, Ny, tx, ty, pitch;
    double dx, dy;
    double rdx, rdy;   // 1/dx and 1/dy, IEEE correctly rounded (host division)
};

// Periodic index map onto 1..N for |offset| ≤ 2N, branch-free (selects only).
__device__ __forceinline__ int wrapi(int i, int N) {
    i = i < 1 ? i + N : i;
    i = i < 1 ? i + N : i;
    i = i > N ? i - N : i;
    i = i > N ? i - N : i;
    return i;
}
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
__device__ __forceinline__ long off(const KG &g, int i, int j) { return (long)(j - 1) * g.pitch + (i - 1); }
// The same offset in 32 bits, and loads and stores through a 32-bit byte offset from a plane's base: the address is
// base + zext(offset), which compiles to the plane's scalar base plus one 32-bit vector offset per access instead of
// 64-bit address arithmetic per load. nk_create rejects grids whose FP64 plane reaches 4 GiB.
__device__ __forceinline__ unsigned uoff(const KG &g, int i, int j) { return (unsigned)((j - 1) * g.pitch + (i - 1)); }
template <typename T> __device__ __forceinline__ T ld32(const T *p, unsigned e) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + e * (unsigned)sizeof(T));
}
template <typename T> __device__ __forceinline__ void st32(T *p, unsigned e, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + e * (unsigned)sizeof(T)) = v;
}
// index maps by location (Center / Face) along x and y
__device__ __forceinline__ int cx(const KG &g, int i) { return g.tx ? clampi(i, 1, g.Nx) : wrapi(i, g.Nx); }
__device__ __forceinline__ int fx(const KG &g, int i) { return g.tx ? clampi(i, 1, g.Nx + 1) : wrapi(i, g.Nx); }
__device__ __forceinline__ int cy(const KG &g, int j) { return g.ty ? clampi(j, 1, g.Ny) : wrapi(j, g.Ny); }
__device__ __forceinline__ int fy(const KG &g, int j) { return g.ty ? clampi(j, 1, g.Ny + 1) : wrapi(j, g.Ny); }
__device__ __forceinline__ int nxf(const KG &g) { return g.Nx + g.tx; }
__device__ __forceinline__ int nyf(const KG &g) { return g.Ny + g.ty; }

// peripheral_node(i, j, k, grid, LX, LY, Center()) (Oceananigans Grids/inactive_node.jl)
__device__ __forceinline__ bool periph_fc(const KG &g, int i, int j) {
    return (g.tx && (i <= 1 || i >= g.Nx + 1)) || (g.ty && (j < 1 || j > g.Ny));
}
__device__ __forceinline__ bool periph_cf(const KG &g, int i, int j) {
    return (g.ty && (j <= 1 || j >= g.Ny + 1)) || (g.tx && (i < 1 || i > g.Nx));
}
__device__ __forceinline__ bool periph_ff(const KG &g, int i, int j) {
    return (g.tx && (i <= 1 || i >= g.Nx + 1)) || (g.ty && (j <= 1 || j >= g.Ny + 1));
}

// Julia Base.max / min / clamp on Float64 (signed zeros as Julia orders them).
__device__ __forceinline__ double jmax(double x, double y) { double d = x - y; return (isnan(x) || isnan(y)) ? d : (signbit(d) ? y : x); }
__device__ __forceinline__ double jmin(double x, double y) { double d = x - y; return (isnan(x) || isnan(y)) ? d : (signbit(d) ? x : y); }
__device__ __forceinline__ double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

// ---------------------------------------------------------------------------------------------------------------
// Activity flags: bit 0 = u point active, bit 1 = v point active (not a wall, not ice free on both sides).
// inactiveᶠᶜᶜ / inactiveᶜᶠᶜ of src/StressBalances/strain_rates.jl:2-3 evaluated once per coefficient update.
// ---------------------------------------------------------------------------------------------------------------
// Branch-free: the flag byte is always loaded from an in-range (clamped or wrapped) address and the range test is a
// select, so every load of a stencil point can be issued before any of them returns.
__device__ __forceinline__ bool uact(const KG &g, const uint8_t *act, int i, int j) {
    const bool in = (!g.tx || (i >= 1 && i <= g.Nx + 1)) && (!g.ty || (j >= 1 && j <= g.Ny));
    const int ii = g.tx ? clampi(i, 1, g.Nx + 1) : wrapi(i, g.Nx);
    const int jj = g.ty ? clampi(j, 1, g.Ny) : wrapi(j, g.Ny);
    return in & (bool)(ld32(act, uoff(g, ii, jj)) & 1);
}
__device__ __forceinline__ bool vact(const KG &g, const uint8_t *act, int i, int j) {
    const bool in = (!g.tx || (i >= 1 && i <= g.Nx)) && (!g.ty || (j >= 1 && j <= g.Ny + 1));
    const int ii = g.tx ? clampi(i, 1, g.Nx) : wrapi(i, g.Nx);
    const int jj = g.ty ? clampi(j, 1, g.Ny + 1) : wrapi(j, g.Ny);
    return in & (bool)((ld32(act, uoff(g, ii, jj)) >> 1) & 1);
}
// interior_u / interior_v with the stiffness mask (src/StressBalances/strain_rates.jl:5-6)
template <typename T> __device__ __forceinline__ T iu(const KG &g, const uint8_t *act, const T *u, int i, int j) {
    const T val = ld32(u, uoff(g, fx(g, i), cy(g, j)));
    return uact(g, act, i, j) ? val : T(0);
}
template <typename T> __device__ __forceinline__ T iv(const KG &g, const uint8_t *act, const T *v, int i, int j) {
    const T val = ld32(v, uoff(g, cx(g, i), fy(g, j)));
    return vact(g, act, i, j) ? val : T(0);
}
// interior_u / interior_v without a stiffness (walls only), as the viscosity sees the velocity
template <typename T> __device__ __forceinline__ T iu_w(const KG &g, const T *u, int i, int j) {
    const T val = ld32(u, uoff(g, fx(g, i), cy(g, j)));
    return periph_fc(g, i, j) ? T(0) : val;
}
template <typename T> __device__ __forceinline__ T iv_w(const KG &g, const T *v, int i, int j) {
    const T val = ld32(v, uoff(g, cx(g, i), fy(g, j)));
    return periph_cf(g, i, j) ? T(0) : val;
}
template <typename T> __device__ __forceinline__ T ldc(const KG &g, const T *f, int i, int j) { return ld32(f, uoff(g, cx(g, i), cy(g, j))); }
template <typename T> __device__ __forceinline__ T ldfc(const KG &g, const T *f, int i, int j) { return ld32(f, uoff(g, fx(g, i), cy(g, j))); }
template <typename T> __device__ __forceinline__ T ldcf(const KG &g, const T *f, int i, int j) { return ld32(f, uoff(g, cx(g, i), fy(g, j))); }
template <typename T> __device__ __forceinline__ T ldff(const KG &g, const T *f, int i, int j) { return ld32(f, uoff(g, fx(g, i), fy(g, j))); }

// Division in the operator and smoother: IEEE a / d in FP64 (the reference's operation order, bit for bit); in the
// FP32 multigrid preconditioner a · rcp(d) with the hardware reciprocal (≤ 1 ulp), which the compiler computes once
// per divisor instead of expanding every quotient into a denormal-safe division sequence.
__device__ __forceinline__ double qd(double a, double d) { return a / d; }
__device__ __forceinline__ float qd(float a, float d) { return a * __builtin_amdgcn_rcpf(d); }
// Division by a grid spacing d with r = RN(1/d) known. FP64: q = a·r, e = q·d − a (exact in one FMA), q − e·r (one FMA).
// With r the correctly rounded reciprocal and no over- or underflow, Markstein's correction step returns the
// correctly rounded quotient, so the result is a / d bit for bit (signed zeros included: for a = ±0, e = +0 and
// q − e·r = q), in three instructions instead of a full division's ten. FP32: a · rcp(d), as qd(float, float).
struct Dv { double d, r; };
__device__ __forceinline__ double qd(double a, Dv v) {
    const double q = a * v.r;
    const double e = __builtin_fma(q, v.d, -a);
    return __builtin_fma(-e, v.r, q);
}
__device__ __forceinline__ float qd(float a, Dv v) { return a * __builtin_amdgcn_rcpf((float)v.d); }
__device__ __forceinline__ Dv dvx(const KG &g) { return Dv{g.dx, g.rdx}; }
__device__ __forceinline__ Dv dvy(const KG &g) { return Dv{g.dy, g.rdy}; }

template <typename T> struct Coef {
    const T *eta, *etaff, *bu, *bv;
    const uint8_t *act;
};

// ---------------------------------------------------------------------------------------------------------------
// The SSA operator at one (i, j): q = Az β p − ∇·T (src/StressBalances/ssa_operator.jl:60-89), with the
// reference's expression order. Centres (i, j), (i−1, j), (i, j−1) and corners (i, j), (i, j+1), (i+1, j).
// ---------------------------------------------------------------------------------------------------------------
// The iterate is read through accessors fu(offset), fv(offset) (plain arrays, or values formed on the fly).
template <typename T, typename FU> __device__ __forceinline__ T iu_f(const KG &g, const uint8_t *act, FU fu, int i, int j) {
    const T val = fu(uoff(g, fx(g, i), cy(g, j)));
    return uact(g, act, i, j) ? val : T(0);
}
template <typename T, typename FV> __device__ __forceinline__ T iv_f(const KG &g, const uint8_t *act, FV fv, int i, int j) {
    const T val = fv(uoff(g, cx(g, i), fy(g, j)));
    return vact(g, act, i, j) ? val : T(0);
}

// The operator's arithmetic with every input behind an accessor: U(a, b), V(a, b) the masked iterate (as iu, iv
// give it), E(a, b) η̄H at a centre (as ldc), FF(a, b) η̄Hᶠᶠ at a corner (as ldff); pu0, pv0, bu0, bv0 and f0 the
// unmasked iterate, drag and activity byte at (i, j). One body for the global-memory and the LDS-tiled kernels.
// The operator's arithmetic on values: the masked iterate at the twelve stencil points (u at (i, j), (i+1, j),
// (i−1, j), (i, j−1), (i, j+1), (i+1, j−1); v at (i, j), (i, j+1), (i−1, j), (i−1, j+1), (i, j−1), (i+1, j)), η̄H at
// the centres (i, j), (i−1, j), (i, j−1), η̄Hᶠᶠ and the peripheral test at the corners (i, j), (i, j+1), (i+1, j),
// the unmasked iterate and drag at (i, j) and the row activities au, av.
template <typename T>
__device__ __forceinline__ void ssa_arith(const KG &g, T u00, T u10, T um0, T u0m, T u0p, T u1m, T v00, T v01, T vm0, T vm1, T v0m,
                                          T v10, T e_c, T e_w, T e_s, T f00, T f01, T f10, bool p00, bool p01, bool p10, T pu0,
                                          T pv0, T bu0, T bv0, bool au, bool av, T wf, T &qu, T &qv) {
    const T dx = (T)g.dx, dy = (T)g.dy, Az = (T)(g.dx * g.dy);
    const Dv DX = dvx(g), DY = dvy(g);
    // strain rates at the three centres
    const T exx_c = qd(u10 - u00, DX), eyy_c = qd(v01 - v00, DY);      // (i, j)
    const T exx_w = qd(u00 - um0, DX), eyy_w = qd(vm1 - vm0, DY);      // (i-1, j)
    const T exx_s = qd(u1m - u0m, DX), eyy_s = qd(v00 - v0m, DY);      // (i, j-1)
    const T txx_c = T(2) * e_c * (T(2) * exx_c + eyy_c), txx_w = T(2) * e_w * (T(2) * exx_w + eyy_w);
    const T tyy_c = T(2) * e_c * (T(2) * eyy_c + exx_c), tyy_s = T(2) * e_s * (T(2) * eyy_s + exx_s);
    // shear strain at the three corners
    T exy_00 = (qd(u00 - u0m, DY) + qd(v00 - vm0, DX)) / T(2);
    T exy_01 = (qd(u0p - u00, DY) + qd(v01 - vm1, DX)) / T(2);
    T exy_10 = (qd(u10 - u1m, DY) + qd(v10 - v00, DX)) / T(2);
    exy_00 = p00 ? wf * exy_00 : exy_00;
    exy_01 = p01 ? wf * exy_01 : exy_01;
    exy_10 = p10 ? wf * exy_10 : exy_10;
    const T txy_00 = T(2) * f00 * exy_00;
    const T txy_01 = T(2) * f01 * exy_01;
    const T txy_10 = T(2) * f10 * exy_10;
    const T divu = (dy * txx_c - dy * txx_w) + (dx * txy_01 - dx * txy_00);
    const T divv = (dx * tyy_c - dx * tyy_s) + (dy * txy_10 - dy * txy_00);
    const T qu_ = Az * bu0 * pu0 - divu;
    const T qv_ = Az * bv0 * pv0 - divv;
    qu = au ? qu_ : pu0;
    qv = av ? qv_ : pv0;
}

template <typename T, typename FU, typename FV, typename FE, typename FF>
__device__ __forceinline__ void ssa_core(const KG &g, int i, int j, FU U, FV V, FE E, FF Ff, T pu0, T pv0, T bu0, T bv0, uint8_t f0,
                                         T wf, T &qu, T &qv) {
    ssa_arith<T>(g, U(i, j), U(i + 1, j), U(i - 1, j), U(i, j - 1), U(i, j + 1), U(i + 1, j - 1), V(i, j), V(i, j + 1), V(i - 1, j),
                 V(i - 1, j + 1), V(i, j - 1), V(i + 1, j), E(i, j), E(i - 1, j), E(i, j - 1), Ff(i, j), Ff(i, j + 1), Ff(i + 1, j),
                 periph_ff(g, i, j), periph_ff(g, i, j + 1), periph_ff(g, i + 1, j), pu0, pv0, bu0, bv0, (f0 & 1) != 0, (f0 & 2) != 0,
                 wf, qu, qv);
}

template <typename T, typename FU, typename FV>
__device__ __forceinline__ void ssa_apply_f(const KG &g, int i, int j, FU fu, FV fv, const Coef<T> &c, T wf, T &qu, T &qv) {
    const uint8_t *a = c.act;
    const unsigned o = uoff(g, i, j);
    ssa_core<T>(g, i, j, [&](int p, int q) { return iu_f<T>(g, a, fu, p, q); }, [&](int p, int q) { return iv_f<T>(g, a, fv, p, q); },
                [&](int p, int q) { return ldc(g, c.eta, p, q); }, [&](int p, int q) { return ldff(g, c.etaff, p, q); }, fu(o), fv(o),
                ld32(c.bu, o), ld32(c.bv, o), ld32(a, o), wf, qu, qv);
}

template <typename T>
__device__ __forceinline__ void ssa_apply(const KG &g, int i, int j, const T *pu, const T *pv, const Coef<T> &c, T wf,
                                          T &qu, T &qv) {
    ssa_apply_f<T>(g, i, j, [pu](unsigned q) { return ld32(pu, q); }, [pv](unsigned q) { return ld32(pv, q); }, c, wf, qu, qv);
}

// Operator diagonal (src/StressBalances/diagonal_preconditioner.jl:20-44)
template <typename T> __device__ __forceinline__ T corner_factor(const KG &g, int i, int j, T wf) { return periph_ff(g, i, j) ? wf : T(1); }
template <typename T> __device__ __forceinline__ T x_nd(const KG &g, const T *eta, int i, int j) { return qd(T(4) * (T)g.dy * ldc(g, eta, i, j), dvx(g)); }
template <typename T> __device__ __forceinline__ T y_nd(const KG &g, const T *eta, int i, int j) { return qd(T(4) * (T)g.dx * ldc(g, eta, i, j), dvy(g)); }
template <typename T> __device__ __forceinline__ T x_sd(const KG &g, const T *ff, int i, int j, T wf) { return qd(corner_factor(g, i, j, wf) * (T)g.dx * ldff(g, ff, i, j), dvy(g)); }
template <typename T> __device__ __forceinline__ T y_sd(const KG &g, const T *ff, int i, int j, T wf) { return qd(corner_factor(g, i, j, wf) * (T)g.dy * ldff(g, ff, i, j), dvx(g)); }

template <typename T> __device__ __forceinline__ T x_diag(const KG &g, const Coef<T> &c, int i, int j, T wf) {
    T d = (T)(g.dx * g.dy) * c.bu[off(g, i, j)] + x_nd(g, c.eta, i, j) + x_nd(g, c.eta, i - 1, j) + x_sd(g, c.etaff, i, j, wf) +
          x_sd(g, c.etaff, i, j + 1, wf);
    return (!(c.act[off(g, i, j)] & 1) || d == T(0)) ? T(1) : d;
}
template <typename T> __device__ __forceinline__ T y_diag(const KG &g, const Coef<T> &c, int i, int j, T wf) {
    T d = (T)(g.dx * g.dy) * c.bv[off(g, i, j)] + y_nd(g, c.eta, i, j) + y_nd(g, c.eta, i, j - 1) + y_sd(g, c.etaff, i, j, wf) +
          y_sd(g, c.etaff, i + 1, j, wf);
    return (!(c.act[off(g, i, j)] & 2) || d == T(0)) ? T(1) : d;
}
// ℓ1 row sums (src/StressBalances/multigrid_preconditioner.jl:195-211)
// with accessors E (η̄H at centres) and Ff (η̄Hᶠᶠ at corners), drag and activity of the row's point
// On values: η̄H at the row's two centres (e0 at (i, j)), η̄Hᶠᶠ and the wall factor at its two corners (f0v, c0 at
// (i, j)), its drag and activity.
template <typename T>
__device__ __forceinline__ T x_l1_vals(const KG &g, T e0, T e1, T f0v, T f1v, T c0, T c1, T bu0, bool active) {
    auto nd = [&](T e) { return qd(T(4) * (T)g.dy * e, dvx(g)); };
    auto sd = [&](T cf, T f) { return qd(cf * (T)g.dx * f, dvy(g)); };
    T s = (T)(g.dx * g.dy) * bu0 + T(2) * (nd(e0) + nd(e1)) + T(4) * (e0 + e1) + T(2) * (sd(c0, f0v) + sd(c1, f1v)) +
          T(2) * (c0 * f0v + c1 * f1v);
    return (!active || s == T(0)) ? T(1) : s;
}
template <typename T>
__device__ __forceinline__ T y_l1_vals(const KG &g, T e0, T e1, T f0v, T f1v, T c0, T c1, T bv0, bool active) {
    auto nd = [&](T e) { return qd(T(4) * (T)g.dx * e, dvy(g)); };
    auto sd = [&](T cf, T f) { return qd(cf * (T)g.dy * f, dvx(g)); };
    T s = (T)(g.dx * g.dy) * bv0 + T(2) * (nd(e0) + nd(e1)) + T(4) * (e0 + e1) + T(2) * (sd(c0, f0v) + sd(c1, f1v)) +
          T(2) * (c0 * f0v + c1 * f1v);
    return (!active || s == T(0)) ? T(1) : s;
}
template <typename T, typename FE, typename FF>
__device__ __forceinline__ T x_l1_f(const KG &g, int i, int j, T wf, FE E, FF Ff, T bu0, uint8_t f0) {
    return x_l1_vals<T>(g, E(i, j), E(i - 1, j), Ff(i, j), Ff(i, j + 1), corner_factor(g, i, j, wf), corner_factor(g, i, j + 1, wf), bu0,
                        (f0 & 1) != 0);
}
template <typename T, typename FE, typename FF>
__device__ __forceinline__ T y_l1_f(const KG &g, int i, int j, T wf, FE E, FF Ff, T bv0, uint8_t f0) {
    return y_l1_vals<T>(g, E(i, j), E(i, j - 1), Ff(i, j), Ff(i + 1, j), corner_factor(g, i, j, wf), corner_factor(g, i + 1, j, wf), bv0,
                        (f0 & 2) != 0);
}
template <typename T> __device__ __forceinline__ T x_l1(const KG &g, const Coef<T> &c, int i, int j, T wf) {
    const unsigned o = uoff(g, i, j);
    return x_l1_f<T>(g, i, j, wf, [&](int p, int q) { return ldc(g, c.eta, p, q); }, [&](int p, int q) { return ldff(g, c.etaff, p, q); },
                     ld32(c.bu, o), ld32(c.act, o));
}
template <typename T> __device__ __forceinline__ T y_l1(const KG &g, const Coef<T> &c, int i, int j, T wf) {
    const unsigned o = uoff(g, i, j);
    return y_l1_f<T>(g, i, j, wf, [&](int p, int q) { return ldc(g, c.eta, p, q); }, [&](int p, int q) { return ldff(g, c.etaff, p, q); },
                     ld32(c.bv, o), ld32(c.act, o));
}

// ---------------------------------------------------------------------------------------------------------------
// Device scalars of one PCG solve and the reduction finishers.
// ---------------------------------------------------------------------------------------------------------------
struct Scal {
    double rho, alpha, beta, pq, rr, rz, res, r0, tol;
    double reltol, abstol;
    int maxiter, fixed, it, done;
    int nonfinite;    // set with done when ‖r‖ is NaN or Inf (a NaN fails every comparison and would stop as converged)
    int done_by_tol;  // Convergence: the last stop met ‖r‖ ≤ tol
    double out[MAXRED];
    double picard_D, picard_X;
    // Mixed precision: the FP32 preconditioner sees r·zs and its z is scaled back by zi = 1/zs. zs = 2^−⌊log₂ ρ⌋ with ρ the
    // residual norm before the iteration's update (‖r₀‖ for a solve's first preconditioning), so r32 is O(1) however
    // small the FP64 residual gets (a warm-started
    // solve to 1e-14 relative takes ‖r‖ to 1e-27 and below, where z32 = r32/ℓ1 would fall into FP32 denormals). A power of
    // two scales every FP32 operation exactly: without underflow, results are the unscaled ones bit for bit.
    double zs, zi;
    // r·z of the latest preconditioning. A finisher folded into k_update_p / k_update_p32 (each block forms β from
    // the slots itself) stores r·z here rather than in rho, which its other blocks still read; the α finisher then
    // moves it into rho. The finishers of k_finish keep rho_new equal to rho.
    double rho_new;
};

enum Fin : int {
    FIN_NONE = 0,     // store totals in out[]
    FIN_ALPHA = 1,    // pq → α
    FIN_BETA_RES = 2, // (r·z, r·r) → β, ρ, res, it, done
    FIN_INIT = 3,     // (r·z, r·r) → ρ, r₀, tol, res, it = 0, done
    FIN_RES = 4,      // r·r → res, it, done (multigrid path: β follows after the V-cycle)
    FIN_BETA = 5,     // r·z → β, ρ
    FIN_PICARD = 6,   // (Δ², X²) → picard_D, picard_X
    FIN_SCALE = 7,    // r·r → zs, zi (the mixed form's first FP32 right-hand side)
};

// zs = 2^−⌊log₂ n⌋ and zi = 1/zs for a norm n (1 for n = 0 or not finite; the exponent kept within ±1000)
__device__ __forceinline__ void scale_of(double n, double &zs, double &zi) {
    const int e = (n > 0.0 && n <= DBL_MAX) ? ilogb(n) : 0;
    const int k = e < -1000 ? 1000 : (e > 1000 ? -1000 : -e);
    zs = ldexp(1.0, k);
    zi = ldexp(1.0, -k);
}
__device__ __forceinline__ void set_scale(Scal *S, double n) { scale_of(n, S->zs, S->zi); }

// The stopping rule of iterate_conjugate_gradient! (src/Solvers/conjugate_gradient.jl:178-200): FixedIterations stops after
// maxiter steps, Convergence when ‖r‖ ≤ tol or at maxiter. A residual norm that is not finite stops the solve and is
// flagged, whatever the control.
__device__ __forceinline__ void set_done(Scal *S) {
    const bool bad = !(S->res <= DBL_MAX);
    S->nonfinite |= bad ? 1 : 0;
    S->done_by_tol = !S->fixed && S->res <= S->tol;
    S->done = bad || (S->fixed ? (S->it >= S->maxiter) : !(S->res > S->tol && S->it < S->maxiter));
}

__device__ __forceinline__ void finish(Scal *S, int fin, const double *t) {
    switch (fin) {
    case FIN_ALPHA:
        S->rho = S->rho_new;
        S->pq = t[0];
        S->alpha = t[0] == 0.0 ? 0.0 : S->rho / t[0];
        set_scale(S, S->res);
        break;
    case FIN_SCALE: set_scale(S, sqrt(t[0])); break;
    case FIN_BETA_RES: {
        double rz = t[0];
        S->beta = S->rho == 0.0 ? 0.0 : rz / S->rho;
        S->rho = rz;
        S->rho_new = rz;
        S->rr = t[1];
        S->res = sqrt(t[1]);
        S->it += 1;
        set_done(S);
        break;
    }
    case FIN_INIT: {
        S->rho = t[0];
        S->rho_new = t[0];
        S->rr = t[1];
        S->r0 = sqrt(t[1]);
        double a = S->reltol * S->r0, b = S->abstol;
        S->tol = jmax(a, b);
        S->res = S->r0;
        S->it = 0;
        S->nonfinite = 0;
        set_done(S);
        break;
    }
    case FIN_RES:
        S->rr = t[0];
        S->res = sqrt(t[0]);
        S->it += 1;
        set_done(S);
        break;
    case FIN_BETA: S->beta = S->rho == 0.0 ? 0.0 : t[0] / S->rho; S->rho = t[0]; S->rho_new = t[0]; break;
    case FIN_PICARD: S->picard_D = sqrt(t[0]); S->picard_X = sqrt(t[1]); break;
    default: break;
    }
    for (int k = 0; k < MAXRED; k++) S->out[k] = t[k];
}

struct Red {
    double *partials;       // MAXRED * capacity slots
    Scal *S;
    int fin;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block-reduce NV values and store the block's sums in its own slots (plain stores; the kernel boundary makes them
// visible). k_finish then sums every block's slots in block order and runs the finisher: no atomics and no
// device-scope fences, which on a multi-XCD part would write back and invalidate the XCD's L2 once per block.
template <int NV> __device__ void reduce_finish(double (&v)[NV], const Red &r) {
    __shared__ double sh[NV][NT / 64];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int nwaves = (blockDim.x * blockDim.y) >> 6;
    const int nblocks = gridDim.x * gridDim.y;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wave] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int w = 0; w < nwaves; w++) s += sh[k][w];
            r.partials[(long)k * nblocks + bid] = s;
        }
    }
}

__device__ __forceinline__ bool solve_done(const Scal *S) { return S != nullptr && S->done; }

// Sum the nv values' block slots of the preceding kernel (nblocks slots each) into tot[] (LDS, MAXRED entries; the
// entries past nv are 0), visible to every thread of the block on return. Each thread first loads its share of every
// value's slots in batches of FIN_BATCH (all loads of a batch in flight together), then adds them in a fixed order,
// so the totals are the same on every run and in every block that forms them (the block must have NT threads). NVM is
// the most values the caller may ask for (each value's total is formed alike whatever NVM is).
constexpr int FIN_BATCH = 8;
template <int NVM>
__device__ __forceinline__ void fold_sum(const double *__restrict__ partials, int nblocks, int nv, double *tot) {
    __shared__ double sh[NVM][NT / 64];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int nth = blockDim.x * blockDim.y, nwaves = nth >> 6;
    double s[NVM];
#pragma unroll
    for (int k = 0; k < NVM; k++) s[k] = 0.0;
    for (int b0 = 0; b0 < nblocks; b0 += nth * FIN_BATCH) {
        double x[NVM][FIN_BATCH];
#pragma unroll
        for (int k = 0; k < NVM; k++)
#pragma unroll
            for (int u = 0; u < FIN_BATCH; u++) {
                const int b = b0 + u * nth + tid;
                x[k][u] = (k < nv && b < nblocks) ? partials[(long)k * nblocks + b] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < NVM; k++)
#pragma unroll
            for (int u = 0; u < FIN_BATCH; u++) s[k] += x[k][u];
    }
#pragma unroll
    for (int k = 0; k < NVM; k++) {
        const double w = wave_sum(s[k]);
        if (lane == 0) sh[k][wave] = w;
    }
    __syncthreads();
    if (tid == 0) {
        double t[MAXRED] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < nv && k < NVM; k++)
            for (int w = 0; w < nwaves; w++) t[k] += sh[k][w];
        for (int k = 0; k < MAXRED; k++) tot[k] = t[k];
    }
    __syncthreads();
}

// Sum NV block partials of the preceding kernel and run the finisher (one block).
__global__ __launch_bounds__(NT) void k_finish(const double *__restrict__ partials, int nblocks, int nv, Scal *S, int fin,
                                               const Scal *gate) {
    if (solve_done(gate)) return;
    __shared__ double tot[MAXRED];
    fold_sum<MAXRED>(partials, nblocks, nv, tot);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double t[MAXRED];
        for (int k = 0; k < MAXRED; k++) t[k] = tot[k];
        finish(S, fin, t);
    }
}

// XCD-aware tile order: the dispatcher hands consecutive workgroup ids to the 8 XCDs in turn, so id → tile is
// remapped to give each XCD one contiguous run of tiles (row-major). Stencil neighbours — the halo rows and columns
// of a tile — are then mostly fetched by the same XCD's L2 instead of by all eight.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_tile(int &bx, int &by) {
    const int n = gridDim.x * gridDim.y;
    const int id = blockIdx.x + blockIdx.y * gridDim.x;
    const int q = n / NXCD, r = n % NXCD;
    const int x = id % NXCD, k = id / NXCD;
    const int t = x * q + (x < r ? x : r) + k;
    bx = t % gridDim.x;
    by = t / gridDim.x;
}

// Two vertically adjacent points per thread (rows j0, j0 + 1) for the finest-level smoother: the stencil rows the
// two points share are loaded once, and half as many waves cover the grid. Grid: blocks_for_rows(Nx, Ny, SROWS).
#ifndef NK_SROWS   // compile-time override for tuning (-DNK_SROWS=4 measured no faster at 1024²)
#define NK_SROWS 2
#endif
constexpr int SROWS = NK_SROWS;
#define NK_IJ2                                                \
    int bx_, by_;                                             \
    xcd_tile(bx_, by_);                                       \
    const int i = bx_ * BX + threadIdx.x + 1;                 \
    const int j0 = by_ * (BY * SROWS) + threadIdx.y * SROWS + 1;

#define NK_IJ                                                 \
    int bx_, by_;                                             \
    xcd_tile(bx_, by_);                                       \
    const int i = bx_ * BX + threadIdx.x + 1;                 \
    const int j = by_ * BY + threadIdx.y + 1;

// ---------------------------------------------------------------------------------------------------------------
// PCG kernels
// ---------------------------------------------------------------------------------------------------------------

// q = A p (+ p·q → α). With b != nullptr writes r = b − A p instead (initial residual, iterative refinement).
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_apply(KG g, const T *__restrict__ pu, const T *__restrict__ pv, Coef<T> c, T wf,
                                              T *__restrict__ qu, T *__restrict__ qv, const T *__restrict__ bu,
                                              const T *__restrict__ bv, Red red, int nred, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    double acc[1] = {0.0};
    if (i <= g.Nx && j <= g.Ny) {
        T a, b;
        ssa_apply(g, i, j, pu, pv, c, wf, a, b);
        const long o = off(g, i, j);
        if (bu) {
            a = T(1) * bu[o] + T(-1) * a;
            b = T(1) * bv[o] + T(-1) * b;
        }
        qu[o] = a;
        qv[o] = b;
        if (nred) acc[0] = (double)pu[o] * (double)a + (double)pv[o] * (double)b;
    }
    if (nred) reduce_finish<1>(acc, red);
}

// The α finisher inside the consumer: every block sums the operator's p·q slots (fold_nb of them) as k_finish would
// and forms α = ρ/(p·q) and the FP32 scale itself; block 0 stores them (and ρ) for the kernels that follow.
__device__ __forceinline__ void fold_alpha(const double *fold, int fold_nb, Scal *S, double &alpha, double &zs) {
    __shared__ double tot[MAXRED];
    fold_sum<1>(fold, fold_nb, 1, tot);
    const double pq = tot[0], rho = S->rho_new;
    alpha = pq == 0.0 ? 0.0 : rho / pq;
    double zi;
    scale_of(S->res, zs, zi);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        S->rho = rho;
        S->pq = pq;
        S->alpha = alpha;
        S->zs = zs;
        S->zi = zi;
        for (int k = 0; k < MAXRED; k++) S->out[k] = tot[k];
    }
}

// Jacobi step of the PCG with the diagonal preconditioner:
//   x += α p; r += (−α) q; z = r / d; (r·z, r·r) → β, ρ, res, done.  With init != 0: z = r / d, p = z, (r·z, r·r).
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_update_diag(KG g, Coef<T> c, T wf, int precond, T *__restrict__ xu, T *__restrict__ xv,
                                                    const T *__restrict__ pu, const T *__restrict__ pv, const T *__restrict__ qu,
                                                    const T *__restrict__ qv, T *__restrict__ ru, T *__restrict__ rv,
                                                    T *__restrict__ zu, T *__restrict__ zv, T *__restrict__ pu_out,
                                                    T *__restrict__ pv_out, int init, Red red, Scal *S, const double *fold,
                                                    int fold_nb) {
    g.tx = TX; g.ty = TY;
    if (solve_done(S)) return;
    double alpha = 0.0, zs;
    if (!init) {   // with fold: α from the operator's p·q slots (the r·z, r·r slots then go to another buffer)
        if (fold) fold_alpha(fold, fold_nb, S, alpha, zs);
        else alpha = S->alpha;
    }
    NK_IJ
    double acc[2] = {0.0, 0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        T r_u, r_v;
        if (init) {
            r_u = ru[o];
            r_v = rv[o];
        } else {
            const T al = (T)alpha;
            xu[o] = T(1) * xu[o] + al * pu[o];
            xv[o] = T(1) * xv[o] + al * pv[o];
            r_u = T(1) * ru[o] + (-al) * qu[o];
            r_v = T(1) * rv[o] + (-al) * qv[o];
            ru[o] = r_u;
            rv[o] = r_v;
        }
        T z_u, z_v;
        if (precond == 1) {
            z_u = r_u / x_diag(g, c, i, j, wf);
            z_v = r_v / y_diag(g, c, i, j, wf);
        } else {
            z_u = T(1) * r_u + T(0) * r_u;
            z_v = T(1) * r_v + T(0) * r_v;
        }
        zu[o] = z_u;
        zv[o] = z_v;
        if (init) {
            pu_out[o] = z_u;
            pv_out[o] = z_v;
        }
        acc[0] = (double)r_u * (double)z_u + (double)r_v * (double)z_v;
        acc[1] = (double)r_u * (double)r_u + (double)r_v * (double)r_v;
    }
    reduce_finish<2>(acc, red);
}

// The β finisher inside the consumer: every block sums the preconditioner's r·z slots (the first value of its
// reduction) and forms β = r·z / ρ with the previous ρ (which no block of the kernel writes); block 0 stores β and r·z
// (as rho_new, moved into rho by α).
__device__ __forceinline__ double fold_beta(const double *fold, int fold_nb, Scal *S) {
    __shared__ double tot[MAXRED];
    fold_sum<1>(fold, fold_nb, 1, tot);
    const double rz = tot[0], beta = S->rho == 0.0 ? 0.0 : rz / S->rho;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        S->beta = beta;
        S->rho_new = rz;
        for (int k = 0; k < MAXRED; k++) S->out[k] = tot[k];
    }
    return beta;
}

// Multigrid PCG: x += α p; r += (−α) q; r·r → res, it, done.
// With r32u: also r32 = (float) (r·zs), the FP32 preconditioner's right-hand side (mixed precision; zs set with α).
// With fold: α comes from the operator's p·q slots (fold_alpha) instead of a k_finish launch; the r·r slots this kernel
// writes must then be another buffer than fold.
template <typename T>
__global__ __launch_bounds__(NT) void k_update_xr(KG g, T *__restrict__ xu, T *__restrict__ xv, const T *__restrict__ pu,
                                                  const T *__restrict__ pv, const T *__restrict__ qu, const T *__restrict__ qv,
                                                  T *__restrict__ ru, T *__restrict__ rv, float *__restrict__ r32u,
                                                  float *__restrict__ r32v, Red red, Scal *S, const double *fold, int fold_nb) {
    if (solve_done(S)) return;
    double alpha, zs;
    if (fold) fold_alpha(fold, fold_nb, S, alpha, zs);
    else {
        alpha = S->alpha;
        zs = S->zs;
    }
    NK_IJ
    double acc[1] = {0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        const T al = (T)alpha;
        xu[o] = T(1) * xu[o] + al * pu[o];
        xv[o] = T(1) * xv[o] + al * pv[o];
        const T a = T(1) * ru[o] + (-al) * qu[o], b = T(1) * rv[o] + (-al) * qv[o];
        ru[o] = a;
        rv[o] = b;
        if (r32u) {
            r32u[o] = (float)((double)a * zs);
            r32v[o] = (float)((double)b * zs);
        }
        acc[0] = (double)a * (double)a + (double)b * (double)b;
    }
    reduce_finish<1>(acc, red);
}

// p = z + β p (or p = z when init). With fold: β from the preconditioner's r·z slots (fold_beta).
template <typename T>
__global__ __launch_bounds__(NT) void k_update_p(KG g, const T *__restrict__ zu, const T *__restrict__ zv, T *__restrict__ pu,
                                                 T *__restrict__ pv, int init, Scal *S, const double *fold, int fold_nb) {
    if (solve_done(S) && !init) return;
    double beta = 0.0;
    if (!init) beta = fold ? fold_beta(fold, fold_nb, S) : S->beta;
    NK_IJ
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        if (init) {
            pu[o] = zu[o];
            pv[o] = zv[o];
        } else {
            const T be = (T)beta;
            pu[o] = T(1) * zu[o] + be * pu[o];
            pv[o] = T(1) * zv[o] + be * pv[o];
        }
    }
}

// k_update_p with the V-cycle's FP32 result: p = z + β p with z = (double) z32 · zi, the value k_sweep_m_z64 would have
// stored as z64, so p is k_update_p's bit for bit without the FP64 copy of z being written and read.
template <typename T>
__global__ __launch_bounds__(NT) void k_update_p32(KG g, const float *__restrict__ zu, const float *__restrict__ zv, T *__restrict__ pu,
                                                   T *__restrict__ pv, Scal *S, const double *fold, int fold_nb) {
    if (solve_done(S)) return;
    const double beta = fold ? fold_beta(fold, fold_nb, S) : S->beta;
    NK_IJ
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        const T be = (T)beta, zi = (T)S->zi;
        pu[o] = T(1) * ((T)zu[o] * zi) + be * pu[o];
        pv[o] = T(1) * ((T)zv[o] * zi) + be * pv[o];
    }
}

// Σ over :xyz of a·b (+ c·d) → finisher
template <typename T>
__global__ __launch_bounds__(NT) void k_dot(KG g, const T *__restrict__ au, const T *__restrict__ av, const T *__restrict__ bu,
                                            const T *__restrict__ bv, const T *__restrict__ cu, const T *__restrict__ cv,
                                            Red red, const Scal *gate) {
    if (solve_done(gate)) return;
    NK_IJ
    double acc[2] = {0.0, 0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        acc[0] = (double)au[o] * (double)bu[o] + (double)av[o] * (double)bv[o];
        if (cu) acc[1] = (double)cu[o] * (double)cu[o] + (double)cv[o] * (double)cv[o];
    }
    reduce_finish<2>(acc, red);
}

// Picard bookkeeping: x⁻ = 1·x⁻ + (−1)·x, then (x⁻·x⁻, x·x) → Δ, X.
__global__ __launch_bounds__(NT) void k_picard_change(KG g, double *__restrict__ mu_, double *__restrict__ mv_,
                                                      const double *__restrict__ xu, const double *__restrict__ xv, Red red) {
    NK_IJ
    double acc[2] = {0.0, 0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        const double a = 1.0 * mu_[o] + (-1.0) * xu[o], b = 1.0 * mv_[o] + (-1.0) * xv[o];
        mu_[o] = a;
        mv_[o] = b;
        acc[0] = a * a + b * b;
        acc[1] = xu[o] * xu[o] + xv[o] * xv[o];
    }
    reduce_finish<2>(acc, red);
}

// ---------------------------------------------------------------------------------------------------------------
// Multigrid kernels
// ---------------------------------------------------------------------------------------------------------------

// Fused ℓ1-Jacobi sweep x_out = x + (b − A x) / s (src/StressBalances/multigrid_preconditioner.jl:165-219).
// One ℓ1-Jacobi sweep x_out = x + (b − A x)/ℓ1 (multigrid_preconditioner.jl:195-219), two rows per thread.
// DOT: Σ b·x_out → finisher (the r·z of the PCG at the finest level). FZ (the last sweep of the mixed form's FP32
// V-cycle): also z64 = (double) x_out · zi (zi undoes the power-of-two scale of the FP32 right-hand side), and the sum
// is Σ r64·z64 with the FP64 residual r64.
template <typename T, bool FZ, bool DOT>
__device__ __forceinline__ void sweep_body(KG g, int i, int j0, const T *__restrict__ xu, const T *__restrict__ xv, const Coef<T> &c,
                                           T wf, const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                           T *__restrict__ ov, const Red &red, const double *__restrict__ r64u,
                                           const double *__restrict__ r64v, double *__restrict__ z64u, double *__restrict__ z64v,
                                           double zi = 1.0) {
    double acc[1] = {0.0};
#pragma unroll
    for (int r = 0; r < SROWS; r++) {
        const int j = j0 + r;
        if (i <= g.Nx && j <= g.Ny) {
            const unsigned o = uoff(g, i, j);
            T qa, qb;
            ssa_apply(g, i, j, xu, xv, c, wf, qa, qb);
            const T x_u = ld32(xu, o), x_v = ld32(xv, o);
            const T b_u = ld32(bu, o), b_v = ld32(bv, o);
            const T nu = x_u + qd(b_u - qa, x_l1(g, c, i, j, wf));
            const T nv = x_v + qd(b_v - qb, y_l1(g, c, i, j, wf));
            st32(ou, o, nu);
            st32(ov, o, nv);
            if (FZ) {
                const double a = (double)nu * zi, b = (double)nv * zi;
                if (z64u) {   // null: the PCG reads the FP32 z (k_update_p32), z64 = (double) z32 is not stored
                    st32(z64u, o, a);
                    st32(z64v, o, b);
                }
                acc[0] += ld32(r64u, o) * a + ld32(r64v, o) * b;
            } else if (DOT) {
                acc[0] += (double)b_u * (double)nu + (double)b_v * (double)nv;
            }
        }
    }
    if (FZ || DOT) reduce_finish<1>(acc, red);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c, T wf,
                                              const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                              T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_body<T, false, false>(g, i, j0, xu, xv, c, wf, bu, bv, ou, ov, Red{}, nullptr, nullptr, nullptr, nullptr);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_dot(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c, T wf,
                                                  const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                                  T *__restrict__ ov, Red red, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_body<T, false, true>(g, i, j0, xu, xv, c, wf, bu, bv, ou, ov, red, nullptr, nullptr, nullptr, nullptr);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_z64(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c, T wf,
                                                  const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                                  T *__restrict__ ov, Red red, const Scal *gate, const double *__restrict__ r64u,
                                                  const double *__restrict__ r64v, double *__restrict__ z64u,
                                                  double *__restrict__ z64v) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_body<T, true, false>(g, i, j0, xu, xv, c, wf, bu, bv, ou, ov, red, r64u, r64v, z64u, z64v, gate ? gate->zi : 1.0);
}

// Neighbour masks of the finest FP32 level: one 16-bit word per point holding every activity and peripheral test the
// operator at (i, j) makes, so a sweep loads one word instead of twelve flag bytes and their range tests.
//   bits 0-5:  uact at (i, j), (i+1, j), (i−1, j), (i, j−1), (i, j+1), (i+1, j−1)
//   bits 6-11: vact at (i, j), (i, j+1), (i−1, j), (i−1, j+1), (i, j−1), (i+1, j)
//   bits 12-14: periph_ff at (i, j), (i, j+1), (i+1, j)
// (the order of ssa_arith's arguments). Rebuilt with the activity flags, once per coefficient update.
template <int TX, int TY>
__global__ __launch_bounds__(NT) void k_nmask(KG g, const uint8_t *__restrict__ act, uint16_t *__restrict__ nm) {
    g.tx = TX; g.ty = TY;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    unsigned m = 0;
    m |= (unsigned)uact(g, act, i, j) << 0;
    m |= (unsigned)uact(g, act, i + 1, j) << 1;
    m |= (unsigned)uact(g, act, i - 1, j) << 2;
    m |= (unsigned)uact(g, act, i, j - 1) << 3;
    m |= (unsigned)uact(g, act, i, j + 1) << 4;
    m |= (unsigned)uact(g, act, i + 1, j - 1) << 5;
    m |= (unsigned)vact(g, act, i, j) << 6;
    m |= (unsigned)vact(g, act, i, j + 1) << 7;
    m |= (unsigned)vact(g, act, i - 1, j) << 8;
    m |= (unsigned)vact(g, act, i - 1, j + 1) << 9;
    m |= (unsigned)vact(g, act, i, j - 1) << 10;
    m |= (unsigned)vact(g, act, i + 1, j) << 11;
    m |= (unsigned)periph_ff(g, i, j) << 12;
    m |= (unsigned)periph_ff(g, i, j + 1) << 13;
    m |= (unsigned)periph_ff(g, i + 1, j) << 14;
    st32(nm, uoff(g, i, j), (uint16_t)m);
}

// x where bit b of m is set, else +0: a one-bit signed field extract (0 or all ones) ANDed with x's bits, the same
// value as (bit ? x : 0) without a compare and a select per value.
__device__ __forceinline__ float keep_if(unsigned m, int b, float x) {
    return __uint_as_float(__float_as_uint(x) & (unsigned)__builtin_amdgcn_sbfe((int)m, b, 1));
}
__device__ __forceinline__ double keep_if(unsigned m, int b, double x) {   // the builtin returns unsigned: widen as int
    return __longlong_as_double(__double_as_longlong(x) & (long long)(int)__builtin_amdgcn_sbfe((int)m, b, 1));
}

// The operator at (i, j) through the neighbour-mask word m of the point (offset o): the same values as ssa_apply —
// the masked iterate, the peripheral corner tests and the row activities from the word — so q is ssa_apply's bit for
// bit, with one 2-byte load in place of twelve flag bytes and their range tests. The centre and corner coefficients,
// the drag and the unmasked iterate at (i, j) are returned for a caller that needs them (the ℓ1 sums of a sweep).
template <typename T> struct MPoint {
    T x_u, x_v, e_c, e_w, e_s, f00, f01, f10, bu0, bv0;
    bool p00, p01, p10, au, av;
};
template <typename T>
__device__ __forceinline__ MPoint<T> ssa_apply_m(const KG &g, int i, int j, unsigned o, unsigned m, const T *__restrict__ xu,
                                                 const T *__restrict__ xv, const Coef<T> &c, T wf, T &qa, T &qb) {
    // x indices of faces and centres beside i, y indices beside j (clamped or wrapped: always in range)
    const int fip = fx(g, i + 1), fim = fx(g, i - 1), cip = cx(g, i + 1), cim = cx(g, i - 1);
    const int cjm = cy(g, j - 1), cjp = cy(g, j + 1), fjm = fy(g, j - 1), fjp = fy(g, j + 1);
    auto mk = [m](int b, T v) { return keep_if(m, b, v); };
    MPoint<T> P;
    P.x_u = ld32(xu, o); P.x_v = ld32(xv, o);
    P.e_c = ld32(c.eta, o); P.e_w = ld32(c.eta, uoff(g, cim, j)); P.e_s = ld32(c.eta, uoff(g, i, cjm));
    P.f00 = ld32(c.etaff, o); P.f01 = ld32(c.etaff, uoff(g, i, fjp)); P.f10 = ld32(c.etaff, uoff(g, fip, j));
    P.bu0 = ld32(c.bu, o); P.bv0 = ld32(c.bv, o);
    P.p00 = (m >> 12) & 1u; P.p01 = (m >> 13) & 1u; P.p10 = (m >> 14) & 1u; P.au = m & 1u; P.av = (m >> 6) & 1u;
    ssa_arith<T>(g, mk(0, P.x_u), mk(1, ld32(xu, uoff(g, fip, j))), mk(2, ld32(xu, uoff(g, fim, j))), mk(3, ld32(xu, uoff(g, i, cjm))),
                 mk(4, ld32(xu, uoff(g, i, cjp))), mk(5, ld32(xu, uoff(g, fip, cjm))), mk(6, P.x_v), mk(7, ld32(xv, uoff(g, i, fjp))),
                 mk(8, ld32(xv, uoff(g, cim, j))), mk(9, ld32(xv, uoff(g, cim, fjp))), mk(10, ld32(xv, uoff(g, i, fjm))),
                 mk(11, ld32(xv, uoff(g, cip, j))), P.e_c, P.e_w, P.e_s, P.f00, P.f01, P.f10, P.p00, P.p01, P.p10, P.x_u, P.x_v,
                 P.bu0, P.bv0, P.au, P.av, wf, qa, qb);
    return P;
}

// Marks loaded values as used here (an empty asm statement reading their registers), so the compiler issues their
// loads before this point and waits for them here instead of sinking them below a following barrier: a tiled kernel
// then has its tile's loads and its own point's loads in flight together, one round trip to memory instead of two.
template <typename T> __device__ __forceinline__ void keep_loaded(const T &v) { asm volatile("" ::"v"(v)); }
template <typename T, int N> __device__ __forceinline__ void keep_loaded(const T (&a)[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) asm volatile("" ::"v"(a[k]));
}

constexpr int ATW = BX + 2, ATH = BY + 2;   // LDS tile of the iterate for a 64 × 4 block (k_apply_mt)
// ssa_apply_m with the iterate from LDS tiles tu, tv whose element [1][1] is point (i0, j0) (see k_apply_mt).
template <typename T>
__device__ __forceinline__ MPoint<T> ssa_apply_mtile(const KG &g, int i, int j, unsigned o, unsigned m, const T (*tu)[ATW],
                                                     const T (*tv)[ATW], int i0, int j0, const Coef<T> &c, T wf, T &qa, T &qb) {
    const int fip = fx(g, i + 1), cim = cx(g, i - 1), cjm = cy(g, j - 1), fjp = fy(g, j + 1);
    auto mk = [m](int b, T v) { return keep_if(m, b, v); };
    auto U = [&](int di, int dj) { return tu[j - j0 + 1 + dj][i - i0 + 1 + di]; };
    auto V = [&](int di, int dj) { return tv[j - j0 + 1 + dj][i - i0 + 1 + di]; };
    MPoint<T> P;
    P.x_u = U(0, 0); P.x_v = V(0, 0);
    P.e_c = ld32(c.eta, o); P.e_w = ld32(c.eta, uoff(g, cim, j)); P.e_s = ld32(c.eta, uoff(g, i, cjm));
    P.f00 = ld32(c.etaff, o); P.f01 = ld32(c.etaff, uoff(g, i, fjp)); P.f10 = ld32(c.etaff, uoff(g, fip, j));
    P.bu0 = ld32(c.bu, o); P.bv0 = ld32(c.bv, o);
    P.p00 = (m >> 12) & 1u; P.p01 = (m >> 13) & 1u; P.p10 = (m >> 14) & 1u; P.au = m & 1u; P.av = (m >> 6) & 1u;
    ssa_arith<T>(g, mk(0, P.x_u), mk(1, U(1, 0)), mk(2, U(-1, 0)), mk(3, U(0, -1)), mk(4, U(0, 1)), mk(5, U(1, -1)), mk(6, P.x_v),
                 mk(7, V(0, 1)), mk(8, V(-1, 0)), mk(9, V(-1, 1)), mk(10, V(0, -1)), mk(11, V(1, 0)), P.e_c, P.e_w, P.e_s, P.f00,
                 P.f01, P.f10, P.p00, P.p01, P.p10, P.x_u, P.x_v, P.bu0, P.bv0, P.au, P.av, wf, qa, qb);
    return P;
}

// k_apply_m with the iterate staged through LDS: a 64 × 4 block reads p on a 66 × 6 tile per component once (each
// element at its clamped or wrapped index, as the stencil would read it), then every stencil value comes from LDS.
// Same values, so the results are k_apply_m's bit for bit.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_apply_mt(KG g, const T *__restrict__ pu, const T *__restrict__ pv, Coef<T> c,
                                                 const uint16_t *__restrict__ nm, T wf, T *__restrict__ qu, T *__restrict__ qv,
                                                 const T *__restrict__ bu, const T *__restrict__ bv, Red red, int nred, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    __shared__ T tu[ATH][ATW], tv[ATH][ATW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * BY + 1;
    const int tid = threadIdx.y * BX + threadIdx.x;
    // tile coordinates beyond the last row or column any point uses are clamped, so every address stays in range
    auto clx = [&](int i) { return i < 0 ? 0 : (i > g.Nx + 1 ? g.Nx + 1 : i); };
    auto cly = [&](int j) { return j < 0 ? 0 : (j > g.Ny + 1 ? g.Ny + 1 : j); };
    for (int e = tid; e < ATH * ATW; e += NT) {
        const int tj = e / ATW, ti = e - tj * ATW;
        const int i = clx(i0 - 1 + ti), j = cly(j0 - 1 + tj);
        tu[tj][ti] = ld32(pu, uoff(g, fx(g, i), cy(g, j)));
        tv[tj][ti] = ld32(pv, uoff(g, cx(g, i), fy(g, j)));
    }
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    double acc[1] = {0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const unsigned o = uoff(g, i, j);
        T a, b;
        const MPoint<T> P = ssa_apply_mtile<T>(g, i, j, o, ld32(nm, o), tu, tv, i0, j0, c, wf, a, b);
        if (bu) {
            a = T(1) * ld32(bu, o) + T(-1) * a;
            b = T(1) * ld32(bv, o) + T(-1) * b;
        }
        st32(qu, o, a);
        st32(qv, o, b);
        if (nred) acc[0] = (double)P.x_u * (double)a + (double)P.x_v * (double)b;
    }
    if (nred) reduce_finish<1>(acc, red);
}

// q = A p (+ p·q), or r = b − A p, through the neighbour masks: k_apply's results bit for bit (finest level).
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_apply_m(KG g, const T *__restrict__ pu, const T *__restrict__ pv, Coef<T> c,
                                                const uint16_t *__restrict__ nm, T wf, T *__restrict__ qu, T *__restrict__ qv,
                                                const T *__restrict__ bu, const T *__restrict__ bv, Red red, int nred, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    double acc[1] = {0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const unsigned o = uoff(g, i, j);
        T a, b;
        const MPoint<T> P = ssa_apply_m<T>(g, i, j, o, ld32(nm, o), pu, pv, c, wf, a, b);
        if (bu) {
            a = T(1) * ld32(bu, o) + T(-1) * a;
            b = T(1) * ld32(bv, o) + T(-1) * b;
        }
        st32(qu, o, a);
        st32(qv, o, b);
        if (nred) acc[0] = (double)P.x_u * (double)a + (double)P.x_v * (double)b;
    }
    if (nred) reduce_finish<1>(acc, red);
}

// The finest-level sweep through the neighbour masks: sweep_body's results bit for bit, the ℓ1 sums formed from the
// coefficients the operator loaded. FZ as in sweep_body. With tu, tv (TILE) the iterate comes from the block's LDS
// tiles, element [1][1] being point (it0, jt0) (load_xtile).
template <typename T, bool FZ, bool DOT, bool TILE = false>
__device__ __forceinline__ void sweep_m_body(KG g, int i, int j0, const T *__restrict__ xu, const T *__restrict__ xv, const Coef<T> &c,
                                             const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                             T *__restrict__ ov, const Red &red, const double *__restrict__ r64u,
                                             const double *__restrict__ r64v, double *__restrict__ z64u, double *__restrict__ z64v,
                                             const T (*tu)[ATW] = nullptr, const T (*tv)[ATW] = nullptr, int it0 = 0, int jt0 = 0,
                                             double zi = 1.0) {
    double acc[1] = {0.0};
#pragma unroll
    for (int r = 0; r < SROWS; r++) {
        const int j = j0 + r;
        if (i <= g.Nx && j <= g.Ny) {
            const unsigned o = uoff(g, i, j);
            T qa, qb;
            const MPoint<T> P = TILE ? ssa_apply_mtile<T>(g, i, j, o, ld32(nm, o), tu, tv, it0, jt0, c, wf, qa, qb)
                                     : ssa_apply_m<T>(g, i, j, o, ld32(nm, o), xu, xv, c, wf, qa, qb);
            const T c00 = P.p00 ? wf : T(1), c01 = P.p01 ? wf : T(1), c10 = P.p10 ? wf : T(1);   // corner_factor
            const T nu = P.x_u + qd(ld32(bu, o) - qa, x_l1_vals<T>(g, P.e_c, P.e_w, P.f00, P.f01, c00, c01, P.bu0, P.au));
            const T nv = P.x_v + qd(ld32(bv, o) - qb, y_l1_vals<T>(g, P.e_c, P.e_s, P.f00, P.f10, c00, c10, P.bv0, P.av));
            st32(ou, o, nu);
            st32(ov, o, nv);
            if (FZ) {
                const double a = (double)nu * zi, b = (double)nv * zi;
                if (z64u) {   // null: the PCG reads the FP32 z (k_update_p32), z64 = (double) z32 is not stored
                    st32(z64u, o, a);
                    st32(z64v, o, b);
                }
                acc[0] += ld32(r64u, o) * a + ld32(r64v, o) * b;
            } else if (DOT) {
                acc[0] += (double)ld32(bu, o) * (double)nu + (double)ld32(bv, o) * (double)nv;
            }
        }
    }
    if (FZ || DOT) reduce_finish<1>(acc, red);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_m(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c,
                                                const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                                T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_m_body<T, false, false>(g, i, j0, xu, xv, c, nm, wf, bu, bv, ou, ov, Red{}, nullptr, nullptr, nullptr, nullptr);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_m_dot(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c,
                                                    const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu,
                                                    const T *__restrict__ bv, T *__restrict__ ou, T *__restrict__ ov, Red red,
                                                    const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_m_body<T, false, true>(g, i, j0, xu, xv, c, nm, wf, bu, bv, ou, ov, red, nullptr, nullptr, nullptr, nullptr);
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_m_z64(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c,
                                                    const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu, const T *__restrict__ bv,
                                                    T *__restrict__ ou, T *__restrict__ ov, Red red, const Scal *gate,
                                                    const double *__restrict__ r64u, const double *__restrict__ r64v,
                                                    double *__restrict__ z64u, double *__restrict__ z64v) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    sweep_m_body<T, true, false>(g, i, j0, xu, xv, c, nm, wf, bu, bv, ou, ov, red, r64u, r64v, z64u, z64v, nullptr, nullptr, 0, 0,
                                 gate ? gate->zi : 1.0);
}

// The iterate of a 64 × (4·SROWS) block on a 66 × (4·SROWS + 2) LDS tile per component, each element read once at its
// clamped or wrapped index as the stencil reads it (the clamp keeps idle tile elements in range).
constexpr int STH = BY * SROWS + 2;
template <typename T>
__device__ __forceinline__ void load_xtile(const KG &g, const T *__restrict__ xu, const T *__restrict__ xv, T (*tu)[ATW], T (*tv)[ATW],
                                           int i0, int j0) {
    const int tid = threadIdx.y * BX + threadIdx.x;
    for (int e = tid; e < STH * ATW; e += NT) {
        const int tj = e / ATW, ti = e - tj * ATW;
        const int i = clampi(i0 - 1 + ti, 0, g.Nx + 1), j = clampi(j0 - 1 + tj, 0, g.Ny + 1);
        tu[tj][ti] = ld32(xu, uoff(g, fx(g, i), cy(g, j)));
        tv[tj][ti] = ld32(xv, uoff(g, cx(g, i), fy(g, j)));
    }
}

// k_sweep_m and k_sweep_m_dot with the iterate from LDS tiles: the same values, results bit for bit.
template <typename T, int TX, int TY, bool FZ, bool DOT>
__device__ __forceinline__ void sweep_mt(KG g, const T *__restrict__ xu, const T *__restrict__ xv, const Coef<T> &c,
                                         const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu, const T *__restrict__ bv,
                                         T *__restrict__ ou, T *__restrict__ ov, const Red &red, const double *__restrict__ r64u,
                                         const double *__restrict__ r64v, double *__restrict__ z64u, double *__restrict__ z64v) {
    __shared__ T tu[STH][ATW], tv[STH][ATW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, jt0 = by_ * (BY * SROWS) + 1;
    load_xtile<T>(g, xu, xv, tu, tv, i0, jt0);
    __syncthreads();
    const int i = i0 + threadIdx.x, j0 = jt0 + threadIdx.y * SROWS;
    sweep_m_body<T, FZ, DOT, true>(g, i, j0, xu, xv, c, nm, wf, bu, bv, ou, ov, red, r64u, r64v, z64u, z64v, tu, tv, i0, jt0);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_mt(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c,
                                                 const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu, const T *__restrict__ bv,
                                                 T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    sweep_mt<T, TX, TY, false, false>(g, xu, xv, c, nm, wf, bu, bv, ou, ov, Red{}, nullptr, nullptr, nullptr, nullptr);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_mt_dot(KG g, const T *__restrict__ xu, const T *__restrict__ xv, Coef<T> c,
                                                     const uint16_t *__restrict__ nm, T wf, const T *__restrict__ bu,
                                                     const T *__restrict__ bv, T *__restrict__ ou, T *__restrict__ ov, Red red,
                                                     const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    sweep_mt<T, TX, TY, false, true>(g, xu, xv, c, nm, wf, bu, bv, ou, ov, red, nullptr, nullptr, nullptr, nullptr);
}

// The first two ℓ1-Jacobi sweeps from x = 0 at the finest level of the FP32 V-cycle, LDS-tiled: x₁ = 0 + b·rcp(ℓ1) is
// pointwise, so a block's 64 × 8 points read x₁ (masked as the operator sees it) on a 66 × 10 tile, η̄H on 65 × 9 and
// η̄Hᶠᶠ on 65 × 9 from LDS; each tile element's clamped or wrapped index, activity test and x₁ are formed once instead
// of once per use. The second sweep is x₂ = x₁ + (b − A x₁)·rcp(ℓ1) with the arithmetic of ssa_core, so the result is
// that of k_sweep_zero followed by k_sweep, bit for bit.
constexpr int TTW = BX + 2, TTH = BY * SROWS + 2;
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_pre_t(KG g, Coef<T> c, T wf, const T *__restrict__ bu, const T *__restrict__ bv,
                                                    const T *__restrict__ rlu, const T *__restrict__ rlv, T *__restrict__ ou,
                                                    T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    __shared__ T su[TTH][TTW], sv[TTH][TTW], se[TTH - 1][TTW - 1], sf[TTH - 1][TTW - 1];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * (BY * SROWS) + 1;
    const int tid = threadIdx.y * BX + threadIdx.x;
    // tile indices beyond the last row or column any point uses are clamped, so every address stays in range
    auto clx = [&](int i) { return i < 0 ? 0 : (i > g.Nx + 1 ? g.Nx + 1 : i); };
    auto cly = [&](int j) { return j < 0 ? 0 : (j > g.Ny + 1 ? g.Ny + 1 : j); };
    auto x1u = [&](unsigned q) { return T(0) + ld32(bu, q) * ld32(rlu, q); };   // as k_sweep_zero forms it
    auto x1v = [&](unsigned q) { return T(0) + ld32(bv, q) * ld32(rlv, q); };
    const int i = i0 + threadIdx.x, jb = j0 + threadIdx.y * SROWS;
    // the tile elements this thread stages (x₁ at the iterate tile, η and η̄ᶠᶠ at theirs), each round unrolled so all
    // loads are in flight together; idle lanes of the last round reread the last element
    constexpr int XT_ITER = (TTH * TTW + NT - 1) / NT, CT_ITER = ((TTH - 1) * (TTW - 1) + NT - 1) / NT;
    T xa[XT_ITER], xb[XT_ITER], ea[CT_ITER], fa[CT_ITER];
#pragma unroll
    for (int k = 0; k < XT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < TTH * TTW ? e0 : TTH * TTW - 1;
        const int tj = e / TTW, ti = e - tj * TTW;
        const int ii = clx(i0 - 1 + ti), jj = cly(j0 - 1 + tj);
        xa[k] = iu_f<T>(g, c.act, x1u, ii, jj);
        xb[k] = iv_f<T>(g, c.act, x1v, ii, jj);
    }
#pragma unroll
    for (int k = 0; k < CT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < (TTH - 1) * (TTW - 1) ? e0 : (TTH - 1) * (TTW - 1) - 1;
        const int tj = e / (TTW - 1), ti = e - tj * (TTW - 1);
        ea[k] = ldc(g, c.eta, clx(i0 - 1 + ti), cly(j0 - 1 + tj));
        fa[k] = ldff(g, c.etaff, clx(i0 + ti), cly(j0 + tj));
    }
    // the points' own values while the tile's loads are in flight
    T b_u[SROWS], b_v[SROWS], r_u[SROWS], r_v[SROWS], be_u[SROWS], be_v[SROWS];
    uint8_t a_c[SROWS];
#pragma unroll
    for (int r = 0; r < SROWS; r++) {
        const int j = jb + r;
        const unsigned o = (i <= g.Nx && j <= g.Ny) ? uoff(g, i, j) : 0u;
        b_u[r] = ld32(bu, o); b_v[r] = ld32(bv, o);
        r_u[r] = ld32(rlu, o); r_v[r] = ld32(rlv, o);
        be_u[r] = ld32(c.bu, o); be_v[r] = ld32(c.bv, o);
        a_c[r] = ld32(c.act, o);
    }
    // idle lanes write the last element's value into it again (no branch)
#pragma unroll
    for (int k = 0; k < XT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < TTH * TTW ? e0 : TTH * TTW - 1;
        const int tj = e / TTW, ti = e - tj * TTW;
        su[tj][ti] = xa[k];
        sv[tj][ti] = xb[k];
    }
#pragma unroll
    for (int k = 0; k < CT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < (TTH - 1) * (TTW - 1) ? e0 : (TTH - 1) * (TTW - 1) - 1;
        const int tj = e / (TTW - 1), ti = e - tj * (TTW - 1);
        se[tj][ti] = ea[k];
        sf[tj][ti] = fa[k];
    }
    keep_loaded(b_u); keep_loaded(b_v); keep_loaded(r_u); keep_loaded(r_v); keep_loaded(be_u); keep_loaded(be_v);
    __syncthreads();
    auto U = [&](int p, int q) { return su[q - j0 + 1][p - i0 + 1]; };
    auto V = [&](int p, int q) { return sv[q - j0 + 1][p - i0 + 1]; };
    auto E = [&](int p, int q) { return se[q - j0 + 1][p - i0 + 1]; };
    auto Ff = [&](int p, int q) { return sf[q - j0][p - i0]; };
#pragma unroll
    for (int r = 0; r < SROWS; r++) {
        const int j = jb + r;
        if (i <= g.Nx && j <= g.Ny) {
            const unsigned o = uoff(g, i, j);
            const T pu0 = T(0) + b_u[r] * r_u[r], pv0 = T(0) + b_v[r] * r_v[r];
            T qa, qb;
            ssa_core<T>(g, i, j, U, V, E, Ff, pu0, pv0, be_u[r], be_v[r], a_c[r], wf, qa, qb);
            st32(ou, o, pu0 + (b_u[r] - qa) * r_u[r]);
            st32(ov, o, pv0 + (b_v[r] - qb) * r_v[r]);
        }
    }
}

// 1/ℓ1 of every row at the finest level (FP32 V-cycle of the mixed form): qd(1, ℓ1) = rcp(ℓ1), the factor every FP32
// sweep applies, stored so the fused first sweeps can form x₁ = b/ℓ1 at a row's neighbours.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_l1_recip(KG g, Coef<T> c, T wf, T *__restrict__ ru, T *__restrict__ rv) {
    g.tx = TX; g.ty = TY;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    ru[o] = qd(T(1), x_l1(g, c, i, j, wf));
    rv[o] = qd(T(1), y_l1(g, c, i, j, wf));
}

// First ℓ1-Jacobi sweep from x = 0: A·0 = 0, so x = 0 + (b − 0)/s exactly as the reference evaluates it.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_sweep_zero(KG g, Coef<T> c, T wf, const T *__restrict__ bu, const T *__restrict__ bv,
                                                   T *__restrict__ ou, T *__restrict__ ov, Red red, int nred, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ2
    double acc[1] = {0.0};
#pragma unroll
    for (int r = 0; r < SROWS; r++) {
        const int j = j0 + r;
        if (i <= g.Nx && j <= g.Ny) {
            const long o = off(g, i, j);
            const T nu = T(0) + qd(bu[o] - T(0), x_l1(g, c, i, j, wf));
            const T nv = T(0) + qd(bv[o] - T(0), y_l1(g, c, i, j, wf));
            ou[o] = nu;
            ov[o] = nv;
            if (nred) acc[0] += (double)bu[o] * (double)nu + (double)bv[o] * (double)nv;
        }
    }
    if (nred) reduce_finish<1>(acc, red);
}

// Thomas algorithm of ℓ1 line Jacobi along y (multigrid_preconditioner.jl:221-259): one thread per x index.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_line_relax(KG g, T *__restrict__ xu, T *__restrict__ xv, const T *__restrict__ bu,
                                                   const T *__restrict__ bv, T *__restrict__ qu, T *__restrict__ qv,
                                                   T *__restrict__ t, Coef<T> c, T wf, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    const int i = blockIdx.x * NT + threadIdx.x + threadIdx.y * BX + 1;
    if (i > g.Nx) return;
    const int Ny = g.Ny;
    for (int comp = 0; comp < 2; comp++) {
        T *x = comp == 0 ? xu : xv;
        const T *b = comp == 0 ? bu : bv;
        T *q = comp == 0 ? qu : qv;
        T cc = 0, tj = 0, d = 0;
        for (int j = 1; j <= Ny; j++) {
            const T s = comp == 0 ? x_l1(g, c, i, j, wf) : y_l1(g, c, i, j, wf);
            T cn;
            if (comp == 0) cn = (uact(g, c.act, i, j) && uact(g, c.act, i, j + 1)) ? x_sd(g, c.etaff, i, j + 1, wf) : T(0);
            else cn = (vact(g, c.act, i, j) && vact(g, c.act, i, j + 1)) ? y_nd(g, c.eta, i, j) : T(0);
            cn = (j == Ny) ? T(0) : cn;
            const T m = s - cc - cn + cc * tj;
            const long o = off(g, i, j);
            d = (b[o] - q[o] + cc * d) / m;
            tj = -cn / m;
            t[o] = tj;
            q[o] = d;
            cc = cn;
        }
        T delta = 0;
        for (int j = Ny; j >= 1; j--) {
            const long o = off(g, i, j);
            delta = q[o] - t[o] * delta;
            x[o] += delta;
        }
    }
}

// restriction_index / restriction_weight / parents (multigrid_preconditioner.jl:264-305)
__device__ __forceinline__ int r_index(int I, int a, int N, int bounded, int face, int h) {
    if (!h) return I;
    if (!bounded) return wrapi(2 * I - 1 + a, N);
    if (face) { int i = 2 * I - 1 + a; return i > 1 ? i : 1; }
    return clampi(2 * I - 1 + a, 1, N);
}
template <typename T> __device__ __forceinline__ T r_weight(int I, int a, int Nc, int bounded, int face, int h) {
    if (!h) return T(1);
    if (face) {
        T w = (a == 0) ? T(1) : T(1) / T(2);
        return (bounded && (2 * I - 1 + a < 1)) ? T(0) : w;
    }
    T w = ((a == 0) | (a == 1)) ? T(3) * T(1) / T(4) : T(1) / T(4);
    if (!bounded) return w;
    const int outside = ((I == 1) & (a == -1)) | ((I == Nc) & (a == 2));
    const int folded = ((I == 1) & (a == 0)) | ((I == Nc) & (a == 1));
    w = w + (folded ? T(1) / T(4) : T(0));
    return outside ? T(0) : w;
}
template <typename T>
__device__ __forceinline__ void parents(int i, int Nc, int bounded, int face, int h, int &I1, int &I2, T &w1, T &w2) {
    if (!h) { I1 = i; I2 = i; w1 = T(1); w2 = T(0); return; }
    const int I = (i + 1) / 2, odd = i & 1;
    if (face) {
        const int up = bounded ? I + 1 : wrapi(I + 1, Nc);
        I1 = I; I2 = odd ? I : up; w1 = odd ? T(1) : T(1) / T(2); w2 = odd ? T(0) : T(1) / T(2);
    } else {
        const int Ip = odd ? I - 1 : I + 1;
        I1 = I; I2 = bounded ? clampi(Ip, 1, Nc) : wrapi(Ip, Nc); w1 = T(3) * T(1) / T(4); w2 = T(1) / T(4);
    }
}

// _restrict! (multigrid_preconditioner.jl:346-352): coarse b = R (fine r with inactive points zeroed)
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_restrict(KG cg, KG fg, int hx, int hy, const uint8_t *__restrict__ cact,
                                                 const uint8_t *__restrict__ fact, const T *__restrict__ ru,
                                                 const T *__restrict__ rv, T *__restrict__ bu, T *__restrict__ bv,
                                                 const Scal *gate) {
    cg.tx = TX; cg.ty = TY;
    fg.tx = TX; fg.ty = TY;
    if (solve_done(gate)) return;
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int I = bx_ * BX + threadIdx.x + 1, J = by_ * BY + threadIdx.y + 1;
    if (I > cg.Nx || J > cg.Ny) return;
    const long o = off(cg, I, J);
    T su = T(0), sv = T(0);
    // Fixed trip counts, so every load of the support is issued before the first is used; offsets outside the
    // support of an un-halved direction read an in-range point with weight 0 (adding +0 leaves each sum unchanged).
    // u: Face along x, Center along y
#pragma unroll
    for (int a = -1; a <= 1; a++)
#pragma unroll
        for (int b = -1; b <= 2; b++) {
            const bool on = (hx || a == 0) && (hy || b == 0);
            const int i = r_index(I, a, fg.Nx, fg.tx, 1, hx), j = r_index(J, b, fg.Ny, fg.ty, 0, hy);
            const T w = r_weight<T>(I, a, cg.Nx, fg.tx, 1, hx) * r_weight<T>(J, b, cg.Ny, fg.ty, 0, hy);
            const T val = iu(fg, fact, ru, i, j);
            if (on) su += w * val;
        }
    // v: Center along x, Face along y
#pragma unroll
    for (int a = -1; a <= 2; a++)
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            const bool on = (hx || a == 0) && (hy || b == 0);
            const int i = r_index(I, a, fg.Nx, fg.tx, 0, hx), j = r_index(J, b, fg.Ny, fg.ty, 1, hy);
            const T w = r_weight<T>(I, a, cg.Nx, fg.tx, 0, hx) * r_weight<T>(J, b, cg.Ny, fg.ty, 1, hy);
            const T val = iv(fg, fact, rv, i, j);
            if (on) sv += w * val;
        }
    bu[o] = (cact[o] & 1) ? su : T(0);
    bv[o] = (cact[o] & 2) ? sv : T(0);
}

// (P e)_u and (P e)_v at fine point (i, j): bilinear weights of the parents, inactive coarse points read as zero.
template <typename T>
__device__ __forceinline__ T prolong_u(const KG &fg, const KG &cg, int hx, int hy, const uint8_t *cact, const T *cu, int i, int j) {
    int I1, I2, J1, J2;
    T wx1, wx2, wy1, wy2;
    parents<T>(i, cg.Nx, fg.tx, 1, hx, I1, I2, wx1, wx2);
    parents<T>(j, cg.Ny, fg.ty, 0, hy, J1, J2, wy1, wy2);
    return wx1 * (wy1 * iu(cg, cact, cu, I1, J1) + wy2 * iu(cg, cact, cu, I1, J2)) +
           wx2 * (wy1 * iu(cg, cact, cu, I2, J1) + wy2 * iu(cg, cact, cu, I2, J2));
}
template <typename T>
__device__ __forceinline__ T prolong_v(const KG &fg, const KG &cg, int hx, int hy, const uint8_t *cact, const T *cv, int i, int j) {
    int I1, I2, J1, J2;
    T wx1, wx2, wy1, wy2;
    parents<T>(i, cg.Nx, fg.tx, 0, hx, I1, I2, wx1, wx2);
    parents<T>(j, cg.Ny, fg.ty, 1, hy, J1, J2, wy1, wy2);
    return wx1 * (wy1 * iv(cg, cact, cv, I1, J1) + wy2 * iv(cg, cact, cv, I1, J2)) +
           wx2 * (wy1 * iv(cg, cact, cv, I2, J1) + wy2 * iv(cg, cact, cv, I2, J2));
}

// _prolongate! (multigrid_preconditioner.jl:354-360): fine x += P (coarse x with inactive points zeroed)
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_prolongate(KG fg, KG cg, int hx, int hy, const uint8_t *__restrict__ fact,
                                                   const uint8_t *__restrict__ cact, const T *__restrict__ cu,
                                                   const T *__restrict__ cv, T *__restrict__ xu, T *__restrict__ xv,
                                                   const Scal *gate) {
    fg.tx = TX; fg.ty = TY;
    cg.tx = TX; cg.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > fg.Nx || j > fg.Ny) return;
    const long o = off(fg, i, j);
    const T pu = prolong_u<T>(fg, cg, hx, hy, cact, cu, i, j);
    const T pv = prolong_v<T>(fg, cg, hx, hy, cact, cv, i, j);
    xu[o] += (fact[o] & 1) ? pu : T(0);
    xv[o] += (fact[o] & 2) ? pv : T(0);
}

// _coarsen_coefficients! (multigrid_preconditioner.jl:331-344), FP64
__global__ __launch_bounds__(NT) void k_coarsen(KG cg, KG fg, int hx, int hy, const double *__restrict__ feta,
                                                const double *__restrict__ fbu, const double *__restrict__ fbv,
                                                double *__restrict__ eta, double *__restrict__ bu, double *__restrict__ bv) {
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int I = bx_ * BX + threadIdx.x + 1, J = by_ * BY + threadIdx.y + 1;
    if (I > cg.Nx || J > cg.Ny) return;
    const int fxf = hx ? 2 : 1, fyf = hy ? 2 : 1;
    const int i = fxf * (I - 1) + 1, j = fyf * (J - 1) + 1;
    const double fAz = fg.dx * fg.dy, cAz = cg.dx * cg.dy;
    double S = 0.0;
    for (int a = 0; a <= fxf - 1; a++)
        for (int b = 0; b <= fyf - 1; b++) S += fAz * ldc(fg, feta, i + a, j + b);
    const long o = off(cg, I, J);
    eta[o] = S / cAz;
    double su = 0.0, sv = 0.0;
    for (int a = (hx ? -1 : 0); a <= (hx ? 1 : 0); a++)
        for (int b = (hy ? -1 : 0); b <= (hy ? 2 : 0); b++) {
            const int ii = r_index(I, a, fg.Nx, fg.tx, 1, hx), jj = r_index(J, b, fg.Ny, fg.ty, 0, hy);
            const double w = r_weight<double>(I, a, cg.Nx, fg.tx, 1, hx) * r_weight<double>(J, b, cg.Ny, fg.ty, 0, hy);
            su += w * (fAz * ldfc(fg, fbu, ii, jj));
        }
    for (int a = (hx ? -1 : 0); a <= (hx ? 2 : 0); a++)
        for (int b = (hy ? -1 : 0); b <= (hy ? 1 : 0); b++) {
            const int ii = r_index(I, a, fg.Nx, fg.tx, 0, hx), jj = r_index(J, b, fg.Ny, fg.ty, 1, hy);
            const double w = r_weight<double>(I, a, cg.Nx, fg.tx, 0, hx) * r_weight<double>(J, b, cg.Ny, fg.ty, 1, hy);
            sv += w * (fAz * ldcf(fg, fbv, ii, jj));
        }
    bu[o] = su / cAz;
    bv[o] = sv / cAz;
}

// η̄Hᶠᶠ = ℑxyᶠᶠ(η̄H) (src/StressBalances/stress_balance_coefficients.jl:26-29) and the activity flags, over the
// (Nx + 1) × (Ny + 1) extent of one level.
__global__ __launch_bounds__(NT) void k_corners_flags(KG g, const double *__restrict__ eta, double *__restrict__ etaff,
                                                      uint8_t *__restrict__ act) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    if (etaff && i <= nxf(g) && j <= nyf(g))
        etaff[o] = ((ldc(g, eta, i - 1, j - 1) + ldc(g, eta, i, j - 1)) / 2.0 + (ldc(g, eta, i - 1, j) + ldc(g, eta, i, j)) / 2.0) / 2.0;
    uint8_t f = 0;
    if (i <= nxf(g) && j <= g.Ny) {
        const bool fr = (ldc(g, eta, i - 1, j) + ldc(g, eta, i, j)) == 0.0;
        if (!(periph_fc(g, i, j) || fr)) f |= 1;
    }
    if (i <= g.Nx && j <= nyf(g)) {
        const bool fr = (ldc(g, eta, i, j - 1) + ldc(g, eta, i, j)) == 0.0;
        if (!(periph_cf(g, i, j) || fr)) f |= 2;
    }
    act[o] = f;
}

// FP64 → FP32 copies of a level's coefficients for the mixed-precision inner solve.
__global__ __launch_bounds__(NT) void k_cast_coeffs(KG g, const double *__restrict__ a, const double *__restrict__ b,
                                                    const double *__restrict__ c, const double *__restrict__ d,
                                                    float *__restrict__ a32, float *__restrict__ b32, float *__restrict__ c32,
                                                    float *__restrict__ d32) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    a32[o] = (float)a[o];
    b32[o] = (float)b[o];
    c32[o] = (float)c[o];
    d32[o] = (float)d[o];
}

// Iterative refinement: r32 = (float) (r64 · zs) and d32 = 0 over :xyz (the inner FP32 solve from zero; zs a power of
// two that brings ‖r‖ to O(1))
__global__ __launch_bounds__(NT) void k_demote(KG g, const double *__restrict__ ru, const double *__restrict__ rv,
                                               float *__restrict__ r32u, float *__restrict__ r32v, float *__restrict__ du,
                                               float *__restrict__ dv, double zs) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    r32u[o] = (float)(ru[o] * zs);
    r32v[o] = (float)(rv[o] * zs);
    du[o] = 0.0f;
    dv[o] = 0.0f;
}

// Mixed precision preconditioning: r32 = (float) (r64 · scale->zs) over :xyz (gated like the PCG kernels)
__global__ __launch_bounds__(NT) void k_demote_r(KG g, const double *__restrict__ ru, const double *__restrict__ rv,
                                                 float *__restrict__ r32u, float *__restrict__ r32v, const Scal *gate,
                                                 const Scal *scale) {
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    const double zs = scale->zs;
    r32u[o] = (float)(ru[o] * zs);
    r32v[o] = (float)(rv[o] * zs);
}

// Mixed precision preconditioning: z64 = (double) z32 · zi and (r·z[, r·r]) → finisher, zi = scale->zi undoing the
// power-of-two scale the FP32 right-hand side carried (k_update_xr, k_demote_r).
__global__ __launch_bounds__(NT) void k_promote_dot(KG g, const float *__restrict__ z32u, const float *__restrict__ z32v,
                                                    double *__restrict__ zu, double *__restrict__ zv, const double *__restrict__ ru,
                                                    const double *__restrict__ rv, Red red, const Scal *gate, const Scal *scale) {
    if (solve_done(gate)) return;
    NK_IJ
    double acc[2] = {0.0, 0.0};
    if (i <= g.Nx && j <= g.Ny) {
        const long o = off(g, i, j);
        const double zi = scale->zi;
        const double a = (double)z32u[o] * zi, b = (double)z32v[o] * zi;
        zu[o] = a;
        zv[o] = b;
        acc[0] = ru[o] * a + rv[o] * b;
        acc[1] = ru[o] * ru[o] + rv[o] * rv[o];
    }
    reduce_finish<2>(acc, red);
}

// Iterative refinement: x64 += (double) d32 · zi over :xyz (zi = 1/zs of k_demote)
__global__ __launch_bounds__(NT) void k_promote_add(KG g, const float *__restrict__ du, const float *__restrict__ dv,
                                                    double *__restrict__ xu, double *__restrict__ xv, double zi) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    xu[o] += (double)du[o] * zi;
    xv[o] += (double)dv[o] * zi;
}

// fill over the whole extent (used for fill!(x, 0) and workspace clearing)
template <typename T>
__global__ __launch_bounds__(NT) void k_fill2(KG g, T *__restrict__ a, T *__restrict__ b, T value, const Scal *gate) {
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long o = off(g, i, j);
    a[o] = value;
    if (b) b[o] = value;
}

// z = 1·x + 0·x over :xyz (copyto!), both components
template <typename T>
__global__ __launch_bounds__(NT) void k_copy2(KG g, const T *__restrict__ xu, const T *__restrict__ xv, T *__restrict__ zu,
                                              T *__restrict__ zv) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    zu[o] = T(1) * xu[o] + T(0) * xu[o];
    zv[o] = T(1) * xv[o] + T(0) * xv[o];
}

// ---------------------------------------------------------------------------------------------------------------
// Coarse tail of the V-cycle in one workgroup: every level from `top` down to the coarsest lives in LDS, and each
// step of v_cycle! (fill, ℓ1 sweep, residual, restriction, prolongation) is one barrier-separated pass of the
// workgroup over a level. The stencil functions are the ones the grid-wide kernels use, so the result is the same
// bit for bit; only the launch latency of ~40 tiny kernels per V-cycle disappears.
// ---------------------------------------------------------------------------------------------------------------
constexpr int TAIL_MAX = 12;
constexpr int TAIL_THREADS = 1024;

template <typename T> struct TailLevel {
    KG lg;                    // LDS geometry (pitch = Nx + 1)
    int gpitch;               // pitch of the level's global arrays
    int hx, hy;               // halved relative to the next finer level
    const T *eta, *etaff, *bu, *bv;
    const uint8_t *act;
    int o_eta, o_etaff, o_bu, o_bv, o_xu, o_xv, o_bu_, o_bv_, o_qu, o_qv, o_act;  // LDS offsets in elements of T / bytes
};

template <typename T> struct TailDesc {
    int n, sweeps, coarsest;
    T wf;
    const T *gbu, *gbv;  // right-hand side of the top tail level (global, top level's pitch)
    T *gxu, *gxv;        // result of the top tail level (global)
    TailLevel<T> lv[TAIL_MAX];
};

template <typename T, int TX, int TY>
__global__ __launch_bounds__(TAIL_THREADS) void k_coarse_cycle(TailDesc<T> d, const Scal *gate) {
    if (solve_done(gate)) return;
    auto geom = [](const KG &k) { KG g = k; g.tx = TX; g.ty = TY; return g; };
    extern __shared__ __align__(16) unsigned char smem[];
    T *sm = reinterpret_cast<T *>(smem);
    const int tid = threadIdx.x;
    const int nth = blockDim.x;
    // load coefficients and flags of every tail level; zero the vectors
    for (int l = 0; l < d.n; l++) {
        const TailLevel<T> &L = d.lv[l];
        const int nx = L.lg.Nx + 1, ny = L.lg.Ny + 1;
        for (int p = tid; p < nx * ny; p += nth) {
            const int i = p % nx, j = p / nx;
            const long go = (long)j * L.gpitch + i;
            const int lo = j * L.lg.pitch + i;
            sm[L.o_eta + lo] = L.eta[go];
            sm[L.o_etaff + lo] = L.etaff[go];
            sm[L.o_bu + lo] = L.bu[go];
            sm[L.o_bv + lo] = L.bv[go];
            smem[L.o_act + lo] = L.act[go];
            sm[L.o_xu + lo] = T(0); sm[L.o_xv + lo] = T(0);
            sm[L.o_qu + lo] = T(0); sm[L.o_qv + lo] = T(0);
            sm[L.o_bu_ + lo] = l == 0 ? d.gbu[go] : T(0);
            sm[L.o_bv_ + lo] = l == 0 ? d.gbv[go] : T(0);
        }
    }
    __syncthreads();
    auto coef = [&](int l) {
        const TailLevel<T> &L = d.lv[l];
        return Coef<T>{sm + L.o_eta, sm + L.o_etaff, sm + L.o_bu, sm + L.o_bv, smem + L.o_act};
    };
    // x = 0 + (b − 0)/s, then sweeps−1 two-phase sweeps q = A x; x += (b − q)/s (Jacobi: all q before any x)
    auto smooth = [&](int l, int sweeps, bool from_zero) {
        const TailLevel<T> &L = d.lv[l];
        const KG g = geom(L.lg);
        const Coef<T> c = coef(l);
        T *xu = sm + L.o_xu, *xv = sm + L.o_xv, *qu = sm + L.o_qu, *qv = sm + L.o_qv;
        const T *bu = sm + L.o_bu_, *bv = sm + L.o_bv_;
        for (int k = 0; k < sweeps; k++) {
            const bool z = from_zero && k == 0;
            if (!z) {
                for (int p = tid; p < g.Nx * g.Ny; p += nth) {
                    const int i = p % g.Nx + 1, j = p / g.Nx + 1;
                    T a, b;
                    ssa_apply(g, i, j, xu, xv, c, d.wf, a, b);
                    const long o = off(g, i, j);
                    qu[o] = a; qv[o] = b;
                }
                __syncthreads();
            }
            for (int p = tid; p < g.Nx * g.Ny; p += nth) {
                const int i = p % g.Nx + 1, j = p / g.Nx + 1;
                const long o = off(g, i, j);
                const T x_u = z ? T(0) : xu[o], x_v = z ? T(0) : xv[o];
                const T q_u = z ? T(0) : qu[o], q_v = z ? T(0) : qv[o];
                xu[o] = x_u + qd(bu[o] - q_u, x_l1(g, c, i, j, d.wf));
                xv[o] = x_v + qd(bv[o] - q_v, y_l1(g, c, i, j, d.wf));
            }
            __syncthreads();
        }
    };
    auto zero_x = [&](int l) {
        const TailLevel<T> &L = d.lv[l];
        const int nx = L.lg.Nx + 1, ny = L.lg.Ny + 1;
        for (int p = tid; p < nx * ny; p += nth) {
            const int lo = (p / nx) * L.lg.pitch + p % nx;
            sm[L.o_xu + lo] = T(0); sm[L.o_xv + lo] = T(0);
        }
        __syncthreads();
    };
    // down leg
    int l = 0;
    for (; l < d.n - 1; l++) {
        const TailLevel<T> &L = d.lv[l], &C = d.lv[l + 1];
        const KG g = geom(L.lg), cg = geom(C.lg);
        const Coef<T> c = coef(l);
        if (d.sweeps == 0) zero_x(l);
        smooth(l, d.sweeps, true);
        T *xu = sm + L.o_xu, *xv = sm + L.o_xv, *qu = sm + L.o_qu, *qv = sm + L.o_qv;
        const T *bu = sm + L.o_bu_, *bv = sm + L.o_bv_;
        for (int p = tid; p < g.Nx * g.Ny; p += nth) {
            const int i = p % g.Nx + 1, j = p / g.Nx + 1;
            T a, b;
            ssa_apply(g, i, j, xu, xv, c, d.wf, a, b);
            const long o = off(g, i, j);
            qu[o] = T(1) * bu[o] + T(-1) * a;
            qv[o] = T(1) * bv[o] + T(-1) * b;
        }
        __syncthreads();
        const uint8_t *cact = smem + C.o_act, *fact = smem + L.o_act;
        T *cbu = sm + C.o_bu_, *cbv = sm + C.o_bv_;
        for (int p = tid; p < cg.Nx * cg.Ny; p += nth) {
            const int I = p % cg.Nx + 1, J = p / cg.Nx + 1;
            T su = T(0), sv = T(0);
            for (int a = (C.hx ? -1 : 0); a <= (C.hx ? 1 : 0); a++)
                for (int b = (C.hy ? -1 : 0); b <= (C.hy ? 2 : 0); b++) {
                    const int i = r_index(I, a, g.Nx, g.tx, 1, C.hx), j = r_index(J, b, g.Ny, g.ty, 0, C.hy);
                    su += r_weight<T>(I, a, cg.Nx, g.tx, 1, C.hx) * r_weight<T>(J, b, cg.Ny, g.ty, 0, C.hy) * iu(g, fact, qu, i, j);
                }
            for (int a = (C.hx ? -1 : 0); a <= (C.hx ? 2 : 0); a++)
                for (int b = (C.hy ? -1 : 0); b <= (C.hy ? 1 : 0); b++) {
                    const int i = r_index(I, a, g.Nx, g.tx, 0, C.hx), j = r_index(J, b, g.Ny, g.ty, 1, C.hy);
                    sv += r_weight<T>(I, a, cg.Nx, g.tx, 0, C.hx) * r_weight<T>(J, b, cg.Ny, g.ty, 1, C.hy) * iv(g, fact, qv, i, j);
                }
            const long o = off(cg, I, J);
            cbu[o] = (cact[o] & 1) ? su : T(0);
            cbv[o] = (cact[o] & 2) ? sv : T(0);
        }
        __syncthreads();
    }
    // coarsest level
    if (d.coarsest == 0) zero_x(l);
    smooth(l, d.coarsest, true);
    // up leg
    for (l = d.n - 2; l >= 0; l--) {
        const TailLevel<T> &L = d.lv[l], &C = d.lv[l + 1];
        const KG g = geom(L.lg), cg = geom(C.lg);
        const uint8_t *cact = smem + C.o_act, *fact = smem + L.o_act;
        const T *cu = sm + C.o_xu, *cv = sm + C.o_xv;
        T *xu = sm + L.o_xu, *xv = sm + L.o_xv;
        for (int p = tid; p < g.Nx * g.Ny; p += nth) {
            const int i = p % g.Nx + 1, j = p / g.Nx + 1;
            int I1, I2, J1, J2;
            T wx1, wx2, wy1, wy2;
            parents<T>(i, cg.Nx, g.tx, 1, C.hx, I1, I2, wx1, wx2);
            parents<T>(j, cg.Ny, g.ty, 0, C.hy, J1, J2, wy1, wy2);
            const T pu = wx1 * (wy1 * iu(cg, cact, cu, I1, J1) + wy2 * iu(cg, cact, cu, I1, J2)) +
                         wx2 * (wy1 * iu(cg, cact, cu, I2, J1) + wy2 * iu(cg, cact, cu, I2, J2));
            parents<T>(i, cg.Nx, g.tx, 0, C.hx, I1, I2, wx1, wx2);
            parents<T>(j, cg.Ny, g.ty, 1, C.hy, J1, J2, wy1, wy2);
            const T pv = wx1 * (wy1 * iv(cg, cact, cv, I1, J1) + wy2 * iv(cg, cact, cv, I1, J2)) +
                         wx2 * (wy1 * iv(cg, cact, cv, I2, J1) + wy2 * iv(cg, cact, cv, I2, J2));
            const long o = off(g, i, j);
            xu[o] += (fact[o] & 1) ? pu : T(0);
            xv[o] += (fact[o] & 2) ? pv : T(0);
        }
        __syncthreads();
        smooth(l, d.sweeps, false);
    }
    // result of the top tail level to global memory
    const TailLevel<T> &L = d.lv[0];
    for (int p = tid; p < L.lg.Nx * L.lg.Ny; p += nth) {
        const int i = p % L.lg.Nx, j = p / L.lg.Nx;
        d.gxu[(long)j * L.gpitch + i] = sm[L.o_xu + j * L.lg.pitch + i];
        d.gxv[(long)j * L.gpitch + i] = sm[L.o_xv + j * L.lg.pitch + i];
    }
}

}  // namespace nk
