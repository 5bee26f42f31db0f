// nk_galerkin.hpp — Galerkin coarse operators A_{l+1} = Pᵀ A_l P for the multigrid preconditioner.
//
// The reference rediscretizes each coarse level from coarsened coefficients (src/StressBalances/
// multigrid_preconditioner.jl:128-141). Where a coarse cell straddles a calving front and an ice-free strip narrower
// than the cell (the one-column ocean of the channel workloads), the coarsened viscosity reaches the domain wall and
// the coarse problem pins the front: the coarse correction then misses the outflow mode and PCG needs hundreds of
// iterations at 1024². The Galerkin product with the V-cycle's own transfers (restriction R = Pᵀ of
// k_restrict / k_prolongate) keeps what the fine operator says about the front.
//
// A coarse operator is an explicit stencil: with the C-grid transfers the product closes on 31 couplings per row.
// Layout: 62 planes per level, plane-major (coefficient k of point o at st[k·plane + o]):
//   u row: [ 0,15) u columns at (dx, dy) ∈ [−1, 1] × [−2, 2], [15,31) v columns at [−2, 1] × [−1, 2]
//   v row: [31,47) u columns at [−1, 2] × [−2, 1], [47,62) v columns at [−2, 2] × [−1, 1]
// (dx, dy) are index offsets of the column from the row point; an inactive row is the identity row. Two more planes
// hold 1/ℓ1 (the inverse absolute row sums) for the ℓ1-Jacobi smoother.
#pragma once
#include "nk_device.hpp"

namespace nk {

constexpr int GS_PLANES = 62;   // stencil planes per level
constexpr int GS_ALLOC = 64;    // + 1/ℓ1 of u rows and of v rows

struct GBlk {
    int base, x0, x1, y0, y1;
};
__host__ __device__ constexpr GBlk gblk(int rc, int cc) {
    return rc == 0 ? (cc == 0 ? GBlk{0, -1, 1, -2, 2} : GBlk{15, -2, 1, -1, 2})
                   : (cc == 0 ? GBlk{31, -1, 2, -2, 1} : GBlk{47, -2, 2, -1, 1});
}
__host__ __device__ constexpr int gidx(const GBlk &b, int dx, int dy) { return b.base + (dx - b.x0) * (b.y1 - b.y0 + 1) + (dy - b.y0); }

// The couplings the finest level's assembled stencil can hold (k_g_assemble: the rediscretized operator's 5-point
// blocks and its 4 cross couplings per row, 9 of the 31 slots); its other slots stay zero from allocation.
__host__ __device__ constexpr bool fin_slot(int rc, int cc, int dx, int dy) {
    return rc == cc ? ((dx == 0 && dy >= -1 && dy <= 1) || (dy == 0 && (dx == -1 || dx == 1)))
                    : (rc == 0 ? (dx >= -1 && dx <= 0 && dy >= 0 && dy <= 1) : (dx >= 0 && dx <= 1 && dy >= -1 && dy <= 0));
}
__host__ __device__ constexpr bool fin_slot_k(int rc, int k) {   // slot k (0..62) of row type rc in pattern order
    for (int cc = 0; cc < 2; cc++) {
        const GBlk b = gblk(rc, cc);
        for (int dx = b.x0; dx <= b.x1; dx++)
            for (int dy = b.y0; dy <= b.y1; dy++)
                if (gidx(b, dx, dy) == k) return fin_slot(rc, cc, dx, dy);
    }
    return false;
}

__device__ __forceinline__ long gplane(const KG &g) { return (long)g.pitch * (long)(g.Ny + 1); }

// Row RC (0 = u, 1 = v) of A x at (i, j); x read through the location's index maps (wrap on Periodic, clamp on
// Bounded, where every out-of-range column has a zero coefficient).
// g_row with the iterate given by accessors fu(offset), fv(offset) instead of arrays
template <typename T, int RC, typename FU, typename FV>
__device__ __forceinline__ T g_row_f(const KG &g, const T *__restrict__ st, long plane, long o, int i, int j, FU fu, FV fv) {
    T q = T(0);
    constexpr GBlk bu = gblk(RC, 0), bv = gblk(RC, 1);
#pragma unroll
    for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
        for (int dy = bu.y0; dy <= bu.y1; dy++)
            q += st[gidx(bu, dx, dy) * plane + o] * fu(off(g, fx(g, i + dx), cy(g, j + dy)));
#pragma unroll
    for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
        for (int dy = bv.y0; dy <= bv.y1; dy++)
            q += st[gidx(bv, dx, dy) * plane + o] * fv(off(g, cx(g, i + dx), fy(g, j + dy)));
    return q;
}

template <typename T, int RC>
__device__ __forceinline__ T g_row(const KG &g, const T *__restrict__ st, long plane, long o, int i, int j,
                                   const T *__restrict__ xu, const T *__restrict__ xv) {
    T q = T(0);
    constexpr GBlk bu = gblk(RC, 0), bv = gblk(RC, 1);
#pragma unroll
    for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
        for (int dy = bu.y0; dy <= bu.y1; dy++)
            q += st[gidx(bu, dx, dy) * plane + o] * xu[off(g, fx(g, i + dx), cy(g, j + dy))];
#pragma unroll
    for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
        for (int dy = bv.y0; dy <= bv.y1; dy++)
            q += st[gidx(bv, dx, dy) * plane + o] * xv[off(g, cx(g, i + dx), fy(g, j + dy))];
    return q;
}

// Symmetric packed copy of a level's stencil, read by the launched sweeps and residuals. The Galerkin operator is
// symmetric, so each coupling pair is kept once: a u row keeps the u columns at or after its own point in (dy, dx)
// order (8 of 15) and all 16 v columns; a v row keeps 8 of its 15 v columns and no u column. A coupling left out is
// read from the column point's row, A(p, q) = A(q, p), so the operator evaluated from the packed planes is exactly
// symmetric (the two rows of a pair of the Pᵀ A P product can differ in the last bit). 32 planes instead of 62. The
// planes carry a GP_HALO-point zero border, so on a Bounded side a coupling to a point outside the level reads 0, as
// the full stencil's does; on a Periodic side the column point's row is found by wrapping.
constexpr int GP_PLANES = 32;
constexpr int GP_HALO = 2;
__host__ __device__ constexpr bool gp_stored(int dx, int dy) { return dy > 0 || (dy == 0 && dx >= 0); }
// slot of a kept offset among block b's kept offsets, in gidx order (dx outer, dy inner)
__host__ __device__ constexpr int gp_rank(const GBlk &b, int dx, int dy) {
    int r = 0;
    for (int a = b.x0; a <= b.x1; a++)
        for (int c = b.y0; c <= b.y1; c++) {
            if (a == dx && c == dy) return r;
            if (gp_stored(a, c)) r++;
        }
    return r;
}
// slots: u row's u columns [0, 8), u row's v columns [8, 24), v row's v columns [24, 32)
__host__ __device__ constexpr int gp_uu(int dx, int dy) { return gp_rank(gblk(0, 0), dx, dy); }
__host__ __device__ constexpr int gp_uv(int dx, int dy) { return 8 + gidx(gblk(0, 1), dx, dy) - gblk(0, 1).base; }
__host__ __device__ constexpr int gp_vv(int dx, int dy) { return 24 + gp_rank(gblk(1, 1), dx, dy); }
__host__ __device__ inline int gp_pitch(const KG &g) { return ((g.Nx + 1 + 2 * GP_HALO) + 63) / 64 * 64; }
__host__ __device__ inline long gp_plane(const KG &g) { return (long)gp_pitch(g) * (long)(g.Ny + 1 + 2 * GP_HALO); }
__device__ __forceinline__ long gp_off(const KG &g, int i, int j) { return (long)(j - 1 + GP_HALO) * gp_pitch(g) + (i - 1 + GP_HALO); }
// the row point of a column (Bounded: as is, the border reads 0; Periodic: wrapped)
__device__ __forceinline__ long gp_nb(const KG &g, int i, int j) { return gp_off(g, g.tx ? i : wrapi(i, g.Nx), g.ty ? j : wrapi(j, g.Ny)); }
// levels that get a packed copy (shape only, so the device-memory estimate can mirror it): no offset aliases there
__host__ __device__ inline bool wants_gpack(const KG &g, int level) { return level >= 1 && g.Nx >= 16 && g.Ny >= 16; }

template <typename T>
__global__ __launch_bounds__(NT) void k_g_pack(KG g, const T *__restrict__ st, T *__restrict__ pk) {
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g), q = gp_off(g, i, j), PP = gp_plane(g);
    constexpr GBlk buu = gblk(0, 0), buv = gblk(0, 1), bvv = gblk(1, 1);
#pragma unroll
    for (int dx = buu.x0; dx <= buu.x1; dx++)
#pragma unroll
        for (int dy = buu.y0; dy <= buu.y1; dy++)
            if (gp_stored(dx, dy)) pk[gp_uu(dx, dy) * PP + q] = st[gidx(buu, dx, dy) * P + o];
#pragma unroll
    for (int dx = buv.x0; dx <= buv.x1; dx++)
#pragma unroll
        for (int dy = buv.y0; dy <= buv.y1; dy++) pk[gp_uv(dx, dy) * PP + q] = st[gidx(buv, dx, dy) * P + o];
#pragma unroll
    for (int dx = bvv.x0; dx <= bvv.x1; dx++)
#pragma unroll
        for (int dy = bvv.y0; dy <= bvv.y1; dy++)
            if (gp_stored(dx, dy)) pk[gp_vv(dx, dy) * PP + q] = st[gidx(bvv, dx, dy) * P + o];
}

// Row RC of A x at (i, j) from the packed planes; x through accessors XU(dx, dy), XV(dx, dy) of the column at offset
// (dx, dy) from the row point.
template <typename T, int RC, typename FU, typename FV>
__device__ __forceinline__ T gp_row_d(const KG &g, const T *__restrict__ pk, int i, int j, FU XU, FV XV) {
    const long PP = gp_plane(g), p = gp_off(g, i, j);
    T q = T(0);
    if constexpr (RC == 0) {
        constexpr GBlk bu = gblk(0, 0), bv = gblk(0, 1);
#pragma unroll
        for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
            for (int dy = bu.y0; dy <= bu.y1; dy++) {
                const T c = gp_stored(dx, dy) ? pk[gp_uu(dx, dy) * PP + p] : pk[gp_uu(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)];
                q += c * XU(dx, dy);
            }
#pragma unroll
        for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
            for (int dy = bv.y0; dy <= bv.y1; dy++) q += pk[gp_uv(dx, dy) * PP + p] * XV(dx, dy);
    } else {
        constexpr GBlk bu = gblk(1, 0), bv = gblk(1, 1);
#pragma unroll
        for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
            for (int dy = bu.y0; dy <= bu.y1; dy++)   // v row to u column = the u row's coupling to this v point
                q += pk[gp_uv(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)] * XU(dx, dy);
#pragma unroll
        for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
            for (int dy = bv.y0; dy <= bv.y1; dy++) {
                const T c = gp_stored(dx, dy) ? pk[gp_vv(dx, dy) * PP + p] : pk[gp_vv(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)];
                q += c * XV(dx, dy);
            }
    }
    return q;
}
// gp_row_d in two halves: the 31 couplings of row RC at (i, j) loaded into registers (gp_coefs), then summed against
// the iterate in gp_row_d's order (gp_row_cf). A tiled kernel issues the coupling loads before it stages its iterate
// tile, so they are in flight together instead of one after the other across the block's barrier.
template <typename T, int RC>
__device__ __forceinline__ void gp_coefs(const KG &g, const T *__restrict__ pk, int i, int j, T (&cf)[31]) {
    const long PP = gp_plane(g), p = gp_off(g, i, j);
    int k = 0;
    if constexpr (RC == 0) {
        constexpr GBlk bu = gblk(0, 0), bv = gblk(0, 1);
#pragma unroll
        for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
            for (int dy = bu.y0; dy <= bu.y1; dy++)
                cf[k++] = gp_stored(dx, dy) ? pk[gp_uu(dx, dy) * PP + p] : pk[gp_uu(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)];
#pragma unroll
        for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
            for (int dy = bv.y0; dy <= bv.y1; dy++) cf[k++] = pk[gp_uv(dx, dy) * PP + p];
    } else {
        constexpr GBlk bu = gblk(1, 0), bv = gblk(1, 1);
#pragma unroll
        for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
            for (int dy = bu.y0; dy <= bu.y1; dy++) cf[k++] = pk[gp_uv(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)];
#pragma unroll
        for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
            for (int dy = bv.y0; dy <= bv.y1; dy++)
                cf[k++] = gp_stored(dx, dy) ? pk[gp_vv(dx, dy) * PP + p] : pk[gp_vv(-dx, -dy) * PP + gp_nb(g, i + dx, j + dy)];
    }
}
template <typename T, int RC, typename FU, typename FV>
__device__ __forceinline__ T gp_row_cf(const T (&cf)[31], FU XU, FV XV) {
    constexpr GBlk bu = gblk(RC, 0), bv = gblk(RC, 1);
    T q = T(0);
    int k = 0;
#pragma unroll
    for (int dx = bu.x0; dx <= bu.x1; dx++)
#pragma unroll
        for (int dy = bu.y0; dy <= bu.y1; dy++) q += cf[k++] * XU(dx, dy);
#pragma unroll
    for (int dx = bv.x0; dx <= bv.x1; dx++)
#pragma unroll
        for (int dy = bv.y0; dy <= bv.y1; dy++) q += cf[k++] * XV(dx, dy);
    return q;
}
// x through accessors fu(offset), fv(offset) as in g_row_f
template <typename T, int RC, typename FU, typename FV>
__device__ __forceinline__ T gp_row_f(const KG &g, const T *__restrict__ pk, int i, int j, FU fu, FV fv) {
    return gp_row_d<T, RC>(g, pk, i, j, [&](int dx, int dy) { return fu(off(g, fx(g, i + dx), cy(g, j + dy))); },
                           [&](int dx, int dy) { return fv(off(g, cx(g, i + dx), fy(g, j + dy))); });
}

// The finest level's operator as a stencil (the ssa_operator! couplings expanded, same masks as ssa_apply).
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_assemble(KG g, Coef<T> c, T wf, T *__restrict__ st) {
    g.tx = TX; g.ty = TY;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    const T dx = (T)g.dx, dy = (T)g.dy, Az = (T)(g.dx * g.dy);
    // only the 18 slots the rediscretized operator can fill are written (zero first); the other 44 stay zero from
    // allocation, and the triple product of the first coarse level reads only these (fin_slot)
#pragma unroll
    for (int k = 0; k < GS_PLANES; k++)
        if (fin_slot_k(k < 31 ? 0 : 1, k)) st[k * P + o] = T(0);
    const T ec = ldc(g, c.eta, i, j), ew = ldc(g, c.eta, i - 1, j), es = ldc(g, c.eta, i, j - 1);
    const T f00 = ldff(g, c.etaff, i, j) * (periph_ff(g, i, j) ? wf : T(1));
    const T f01 = ldff(g, c.etaff, i, j + 1) * (periph_ff(g, i, j + 1) ? wf : T(1));
    const T f10 = ldff(g, c.etaff, i + 1, j) * (periph_ff(g, i + 1, j) ? wf : T(1));
    const uint8_t *a = c.act;
    constexpr GBlk UU = gblk(0, 0), UV = gblk(0, 1), VU = gblk(1, 0), VV = gblk(1, 1);
    if (uact(g, a, i, j)) {
        st[gidx(UU, 0, 0) * P + o] = Az * c.bu[o] + T(4) * dy * ec / dx + T(4) * dy * ew / dx + dx * f01 / dy + dx * f00 / dy;
        if (uact(g, a, i + 1, j)) st[gidx(UU, 1, 0) * P + o] = -T(4) * dy * ec / dx;
        if (uact(g, a, i - 1, j)) st[gidx(UU, -1, 0) * P + o] = -T(4) * dy * ew / dx;
        if (uact(g, a, i, j + 1)) st[gidx(UU, 0, 1) * P + o] = -dx * f01 / dy;
        if (uact(g, a, i, j - 1)) st[gidx(UU, 0, -1) * P + o] = -dx * f00 / dy;
        if (vact(g, a, i, j + 1)) st[gidx(UV, 0, 1) * P + o] = -T(2) * ec - f01;
        if (vact(g, a, i, j)) st[gidx(UV, 0, 0) * P + o] = T(2) * ec + f00;
        if (vact(g, a, i - 1, j + 1)) st[gidx(UV, -1, 1) * P + o] = T(2) * ew + f01;
        if (vact(g, a, i - 1, j)) st[gidx(UV, -1, 0) * P + o] = -T(2) * ew - f00;
    } else {
        st[gidx(UU, 0, 0) * P + o] = T(1);
    }
    if (vact(g, a, i, j)) {
        st[gidx(VV, 0, 0) * P + o] = Az * c.bv[o] + T(4) * dx * ec / dy + T(4) * dx * es / dy + dy * f10 / dx + dy * f00 / dx;
        if (vact(g, a, i, j + 1)) st[gidx(VV, 0, 1) * P + o] = -T(4) * dx * ec / dy;
        if (vact(g, a, i, j - 1)) st[gidx(VV, 0, -1) * P + o] = -T(4) * dx * es / dy;
        if (vact(g, a, i + 1, j)) st[gidx(VV, 1, 0) * P + o] = -dy * f10 / dx;
        if (vact(g, a, i - 1, j)) st[gidx(VV, -1, 0) * P + o] = -dy * f00 / dx;
        if (uact(g, a, i + 1, j)) st[gidx(VU, 1, 0) * P + o] = -T(2) * ec - f10;
        if (uact(g, a, i, j)) st[gidx(VU, 0, 0) * P + o] = T(2) * ec + f00;
        if (uact(g, a, i + 1, j - 1)) st[gidx(VU, 1, -1) * P + o] = T(2) * es + f10;
        if (uact(g, a, i, j - 1)) st[gidx(VU, 0, -1) * P + o] = -T(2) * es - f00;
    } else {
        st[gidx(VV, 0, 0) * P + o] = T(1);
    }
}

// Offset d of a coarse column from the row along one direction, as the pattern range [lo, hi] stores it (periodic
// directions fold modulo N). Returns false when no representative lies in the range.
__device__ __forceinline__ bool g_fold(int &d, int lo, int hi, int N, int periodic) {
    if (periodic) {
        while (d < lo) d += N;
        while (d > hi) d -= N;
    }
    return d >= lo && d <= hi;
}

// Activity of the coarse columns within ±2 of (I, J) as bits of the mask g_rap_point tests: bit (dx + 2)·5 + (dy + 2)
// for u, + 25 for v; one offset (dx, dy) = (k / 5 − 2, k % 5 − 2) of the 25.
__device__ __forceinline__ uint64_t g_cmask_bits(const KG &cg, const uint8_t *cact, int I, int J, int k) {
    const int dx = k / 5 - 2, dy = k % 5 - 2;
    int ic = I + dx, jc = J + dy;
    if (!cg.tx) ic = wrapi(ic, cg.Nx);
    if (!cg.ty) jc = wrapi(jc, cg.Ny);
    return ((uint64_t)uact(cg, cact, ic, jc) << k) | ((uint64_t)vact(cg, cact, ic, jc) << (25 + k));
}

// Contribution of one fine point f = support point (a, b) of the restriction of coarse row RC at (I, J):
// w_f · Σ over the fine couplings f′ of f of A_l(f, f′) · P(f′, C), over the coarse parents C of f′ (coarse-active,
// from cmask: g_cmask_bits over the 25 offsets), added into acc[0..31) (the row's couplings in pattern order).
// FIN: the fine level is the finest, whose stencil holds only the fin_slot couplings (the others are exact zeros,
// which this loop skips anyway), so only those are loaded and visited.
// half ≥ 0: only the fine couplings to component half (columns of that block of the pattern, whose coarse targets are
// that block's slots, disjoint from the other's), so two lanes can share a support point without changing any slot's
// order of additions.
template <typename T, int RC, bool FIN>
__device__ __forceinline__ void g_rap_point(const KG &cg, const KG &fg, int hx, int hy, uint64_t cmask, const uint8_t *fact,
                                            const T *__restrict__ fst, long fplane, int I, int J, int a, int b, T *acc, int *err,
                                            int half) {
    constexpr int base = gblk(RC, 0).base;
    const int fr = RC == 0 ? 1 : 0;   // restriction: u is Face along x, v is Face along y
    const int i = r_index(I, a, fg.Nx, fg.tx, fr, hx), j = r_index(J, b, fg.Ny, fg.ty, 1 - fr, hy);
    const T w = r_weight<T>(I, a, cg.Nx, fg.tx, fr, hx) * r_weight<T>(J, b, cg.Ny, fg.ty, 1 - fr, hy);
    if (w == T(0)) return;
    if (!(RC == 0 ? uact(fg, fact, i, j) : vact(fg, fact, i, j))) return;
    const long fo = off(fg, i, j);
    constexpr int nu = gblk(RC, 1).base - base;   // couplings to u columns; the rest are to v columns
    T av[31];
#pragma unroll
    for (int k = 0; k < 31; k++)
        av[k] = ((!FIN || fin_slot_k(RC, base + k)) && (half < 0 || (k < nu) == (half == 0))) ? fst[(base + k) * fplane + fo] : T(0);
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
        if (half >= 0 && cc != half) continue;
        constexpr GBlk B0 = gblk(RC, 0), B1 = gblk(RC, 1);
        const GBlk fb = cc == 0 ? B0 : B1;
#pragma unroll
        for (int dx = fb.x0; dx <= fb.x1; dx++)
#pragma unroll
            for (int dy = fb.y0; dy <= fb.y1; dy++) {
                if (FIN && !fin_slot(RC, cc, dx, dy)) continue;
                const T c = av[gidx(fb, dx, dy) - base];
                if (c == T(0)) continue;
                int i2 = i + dx, j2 = j + dy;
                if (!fg.tx) i2 = wrapi(i2, fg.Nx);
                if (!fg.ty) j2 = wrapi(j2, fg.Ny);
                int I1, I2, J1, J2;
                T wx1, wx2, wy1, wy2;
                parents<T>(i2, cg.Nx, fg.tx, cc == 0 ? 1 : 0, hx, I1, I2, wx1, wx2);
                parents<T>(j2, cg.Ny, fg.ty, cc == 0 ? 0 : 1, hy, J1, J2, wy1, wy2);
                const int Is[2] = {I1, I2}, Js[2] = {J1, J2};
                const T wxs[2] = {wx1, wx2}, wys[2] = {wy1, wy2};
#pragma unroll
                for (int p = 0; p < 2; p++)
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        const T wp = wxs[p] * wys[q];
                        if (wp == T(0)) continue;
                        int ddx = Is[p] - I, ddy = Js[q] - J;
                        int mx = ddx, my = ddy;
                        if (!g_fold(mx, -2, 2, cg.Nx, !cg.tx) || !g_fold(my, -2, 2, cg.Ny, !cg.ty)) {
                            atomicAdd(err, 1);
                            continue;
                        }
                        if (!((cmask >> (cc * 25 + (mx + 2) * 5 + (my + 2))) & 1)) continue;
                        // the coarse row couples to component cc through the same block of the pattern
                        if (!g_fold(ddx, fb.x0, fb.x1, cg.Nx, !cg.tx) || !g_fold(ddy, fb.y0, fb.y1, cg.Ny, !cg.ty)) {
                            atomicAdd(err, 1);
                            continue;
                        }
                        acc[gidx(fb, ddx, ddy) - base] += w * c * wp;
                    }
            }
    }
}

// A_{l+1} = Pᵀ A_l P and 1/ℓ1 of its rows; inactive coarse rows become identity rows. One 256-thread block per
// GRAP_POINTS coarse points (k_g_rap): 16 row slots (u and v of each point) × 16 lanes, lane m < 12 handling support
// point m of the row's restriction into its own 31 partial sums in LDS. k_g_rap_split: half as many points per block
// and each support point's couplings divided between two lanes by column component (threads 0–127 the u columns,
// 128–255 the v columns, so a wave takes one half only): the halves fill disjoint slots, every slot's additions keep
// their order, and each lane's chain of loads and LDS updates is about half as long. The partials are then summed over
// the support points in order (deterministic). One-dimensional grid of ⌈Nx·Ny / points per block⌉ blocks.
constexpr int GRAP_POINTS = 8;
template <typename T, int TX, int TY, int SPLIT>
__device__ __forceinline__ void g_rap_block(KG cg, KG fg, int hx, int hy, const uint8_t *__restrict__ cact,
                                            const uint8_t *__restrict__ fact, const T *__restrict__ fst, T *__restrict__ cst,
                                            int *err, int fin, T *__restrict__ ccst) {
    cg.tx = TX; cg.ty = TY;
    fg.tx = TX; fg.ty = TY;
    constexpr int npb = SPLIT ? GRAP_POINTS / 2 : GRAP_POINTS;   // coarse points per block
    __shared__ T part[2 * npb][12][31];
    __shared__ T absv[2 * npb][31];
    __shared__ unsigned long long cmask[npb];
    __shared__ int sact[2 * npb];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int half = SPLIT ? (tid >> 7) : -1;
    const int slot = SPLIT ? ((tid & 127) >> 4) : (tid >> 4), m = tid & 15, rc = slot & 1;
    const int npts = cg.Nx * cg.Ny;
    const int q = blockIdx.x * npb + (slot >> 1);
    const int I = q % cg.Nx + 1, J = q / cg.Nx + 1;
    const bool active = q < npts && (rc == 0 ? uact(cg, cact, I, J) : vact(cg, cact, I, J));
    if (tid < npb) cmask[tid] = 0ull;
    if (m == 0 && half <= 0) sact[slot] = active ? 1 : 0;
    __syncthreads();
    {   // the coarse activity around each of the block's points, formed once per point (32 threads each, 25 offsets)
        const int p = tid / 32, k = tid % 32, qp = blockIdx.x * npb + p;
        if (p < npb && k < 25 && qp < npts)
            atomicOr(&cmask[p], (unsigned long long)g_cmask_bits(cg, cact, qp % cg.Nx + 1, qp / cg.Nx + 1, k));
    }
    __syncthreads();
    if (m < 12) {
        T *acc = part[slot][m];
        const int nu = gblk(rc, 1).base - gblk(rc, 0).base;
        const int k0 = half == 1 ? nu : 0, k1 = half == 0 ? nu : 31;
        for (int k = k0; k < k1; k++) acc[k] = T(0);
        if (active) {
            const int a = rc == 0 ? m / 4 - 1 : m / 3 - 1, b = rc == 0 ? m % 4 - 1 : m % 3 - 1;
            const bool on = (hx || a == 0) && (hy || b == 0);
            if (on) {
                const long fP = gplane(fg);
                const uint64_t cm = cmask[slot >> 1];
                if (fin) {
                    if (rc == 0) g_rap_point<T, 0, true>(cg, fg, hx, hy, cm, fact, fst, fP, I, J, a, b, acc, err, half);
                    else g_rap_point<T, 1, true>(cg, fg, hx, hy, cm, fact, fst, fP, I, J, a, b, acc, err, half);
                } else {
                    if (rc == 0) g_rap_point<T, 0, false>(cg, fg, hx, hy, cm, fact, fst, fP, I, J, a, b, acc, err, half);
                    else g_rap_point<T, 1, false>(cg, fg, hx, hy, cm, fact, fst, fP, I, J, a, b, acc, err, half);
                }
            }
        }
    }
    __syncthreads();
    // each coupling k of each row: the partials summed over the 12 support points in support order. With ccst the
    // same values also go to the level's compact copy (k_g_compact's layout; its padding points stay zero from
    // allocation), which the LDS tail reads.
    const long P = gplane(cg), NP = (long)(cg.Nx + 1) * (cg.Ny + 1);
    for (int e = tid; e < 2 * npb * 31; e += NT) {
        const int ss = e / 31, k = e % 31, rs = ss & 1, qs = blockIdx.x * npb + (ss >> 1);
        if (qs >= npts) continue;
        const int base = gblk(rs, 0).base, diag = gidx(gblk(rs, rs), 0, 0) - base;
        T v = T(0);
        if (sact[ss]) {
            for (int mm = 0; mm < 12; mm++) v += part[ss][mm][k];
        } else {
            v = k == diag ? T(1) : T(0);
        }
        const int Is = qs % cg.Nx + 1, Js = qs / cg.Nx + 1;
        cst[(base + k) * P + off(cg, Is, Js)] = v;
        if (ccst) ccst[(base + k) * NP + (long)(Js - 1) * (cg.Nx + 1) + (Is - 1)] = v;
        absv[ss][k] = fabs(v);
    }
    __syncthreads();
    if (tid < 2 * npb) {   // 1/ℓ1 of each row, its |couplings| added in pattern order
        const int ss = tid, rs = ss & 1, qs = blockIdx.x * npb + (ss >> 1);
        if (qs < npts) {
            T l1 = T(0);
            for (int k = 0; k < 31; k++) l1 += absv[ss][k];
            const int Is = qs % cg.Nx + 1, Js = qs / cg.Nx + 1;
            const T r = l1 > T(0) ? T(1) / l1 : T(1);
            cst[(GS_PLANES + rs) * P + off(cg, Is, Js)] = r;
            if (ccst) ccst[(GS_PLANES + rs) * NP + (long)(Js - 1) * (cg.Nx + 1) + (Is - 1)] = r;
        }
    }
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_rap(KG cg, KG fg, int hx, int hy, const uint8_t *__restrict__ cact,
                                              const uint8_t *__restrict__ fact, const T *__restrict__ fst, T *__restrict__ cst,
                                              int *err, int fin, T *__restrict__ ccst) {
    g_rap_block<T, TX, TY, 0>(cg, fg, hx, hy, cact, fact, fst, cst, err, fin, ccst);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_rap_split(KG cg, KG fg, int hx, int hy, const uint8_t *__restrict__ cact,
                                                    const uint8_t *__restrict__ fact, const T *__restrict__ fst,
                                                    T *__restrict__ cst, int *err, int fin, T *__restrict__ ccst) {
    g_rap_block<T, TX, TY, 1>(cg, fg, hx, hy, cact, fact, fst, cst, err, fin, ccst);
}

// ℓ1-Jacobi sweep with a stencil operator: x_out = x + (b − A x)/ℓ1, or b/ℓ1 from a zero x.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_sweep(KG g, const T *__restrict__ st, const T *__restrict__ xu, const T *__restrict__ xv,
                                                const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                                T *__restrict__ ov, int from_zero, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    const T su = st[GS_PLANES * P + o], sv = st[(GS_PLANES + 1) * P + o];
    if (from_zero) {
        ou[o] = bu[o] * su;
        ov[o] = bv[o] * sv;
        return;
    }
    const T qu = g_row<T, 0>(g, st, P, o, i, j, xu, xv);
    const T qv = g_row<T, 1>(g, st, P, o, i, j, xu, xv);
    ou[o] = xu[o] + (bu[o] - qu) * su;
    ov[o] = xv[o] + (bv[o] - qv) * sv;
}

// The first two ℓ1-Jacobi sweeps from a zero x in one pass: the first sweep is pointwise (x₁ = b/ℓ1), so the second,
// x₂ = x₁ + (b − A x₁)/ℓ1, reads x₁ at the stencil's neighbours as b·(1/ℓ1) there. Same products and order as the two
// separate sweeps, so the result is the same bit for bit.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_sweep_pre(KG g, const T *__restrict__ st, const T *__restrict__ bu, const T *__restrict__ bv,
                                                    T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    const T *su_p = st + GS_PLANES * P, *sv_p = st + (GS_PLANES + 1) * P;
    const T su = su_p[o], sv = sv_p[o];
    auto x1u = [&](long q) { return bu[q] * su_p[q]; };
    auto x1v = [&](long q) { return bv[q] * sv_p[q]; };
    const T qu = g_row_f<T, 0>(g, st, P, o, i, j, x1u, x1v);
    const T qv = g_row_f<T, 1>(g, st, P, o, i, j, x1u, x1v);
    ou[o] = bu[o] * su + (bu[o] - qu) * su;
    ov[o] = bv[o] * sv + (bv[o] - qv) * sv;
}

// r = b − A x with a stencil operator
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_g_residual(KG g, const T *__restrict__ st, const T *__restrict__ xu, const T *__restrict__ xv,
                                                   const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ru,
                                                   T *__restrict__ rv, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    ru[o] = bu[o] - g_row<T, 0>(g, st, P, o, i, j, xu, xv);
    rv[o] = bv[o] - g_row<T, 1>(g, st, P, o, i, j, xu, xv);
}

// The launched sweeps and residual with the symmetric packed planes pk (1/ℓ1 still from st's last two planes).
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                                 const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv,
                                                 T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    const T su = st[GS_PLANES * P + o], sv = st[(GS_PLANES + 1) * P + o];
    auto fu = [&](long q) { return xu[q]; };
    auto fv = [&](long q) { return xv[q]; };
    const T qu = gp_row_f<T, 0>(g, pk, i, j, fu, fv);
    const T qv = gp_row_f<T, 1>(g, pk, i, j, fu, fv);
    ou[o] = xu[o] + (bu[o] - qu) * su;
    ov[o] = xv[o] + (bv[o] - qv) * sv;
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_pre(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ bu,
                                                     const T *__restrict__ bv, T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j), P = gplane(g);
    const T *su_p = st + GS_PLANES * P, *sv_p = st + (GS_PLANES + 1) * P;
    const T su = su_p[o], sv = sv_p[o];
    auto x1u = [&](long q) { return bu[q] * su_p[q]; };
    auto x1v = [&](long q) { return bv[q] * sv_p[q]; };
    const T qu = gp_row_f<T, 0>(g, pk, i, j, x1u, x1v);
    const T qv = gp_row_f<T, 1>(g, pk, i, j, x1u, x1v);
    ou[o] = bu[o] * su + (bu[o] - qu) * su;
    ov[o] = bv[o] * sv + (bv[o] - qv) * sv;
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_residual(KG g, const T *__restrict__ pk, const T *__restrict__ xu, const T *__restrict__ xv,
                                                    const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ru,
                                                    T *__restrict__ rv, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    NK_IJ
    if (i > g.Nx || j > g.Ny) return;
    const long o = off(g, i, j);
    auto fu = [&](long q) { return xu[q]; };
    auto fv = [&](long q) { return xv[q]; };
    ru[o] = bu[o] - gp_row_f<T, 0>(g, pk, i, j, fu, fv);
    rv[o] = bv[o] - gp_row_f<T, 1>(g, pk, i, j, fu, fv);
}

// The packed sweep and residual with x staged through LDS: a 64 × 4 block reads x on a 68 × 8 tile per component (the
// stencil reaches two points each way), each element once at its clamped or wrapped index as gp_row_f reads it. Same
// values and order, so the results are k_gp_sweep's and k_gp_residual's bit for bit.
constexpr int GTW = BX + 4, GTH = BY + 4;

// The coarse correction of k_prolongate (coarse grid, refinement flags, activity flags, coarse e): with it the tile
// load forms x + P e itself.
template <typename T> struct ProSrc {
    KG cg;
    int hx, hy;
    const uint8_t *fact, *cact;
    const T *cu, *cv;
};
// The tile is read in two phases: every element a thread stages is loaded into registers first (a fixed number of
// rounds, unrolled, so all of a thread's loads are in flight together instead of one round trip per element), then
// written to LDS; a kernel issues the loads of its own point's operands between the two, so they overlap too.
constexpr int GT_ITER = (GTH * GTW + NT - 1) / NT;
template <typename T>
__device__ __forceinline__ void fetch_gtile(const KG &g, const T *__restrict__ xu, const T *__restrict__ xv, int i0, int j0,
                                            T (&va)[GT_ITER], T (&vb)[GT_ITER]) {
    const int tid = threadIdx.y * BX + threadIdx.x;
#pragma unroll
    for (int k = 0; k < GT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < GTH * GTW ? e0 : GTH * GTW - 1;   // idle lanes reread the last element
        const int tj = e / GTW, ti = e - tj * GTW;
        // every column a real point reads lies in [−1, N + 2]; beyond it the clamp only keeps idle elements in range
        const int i = clampi(i0 - 2 + ti, -1, g.Nx + 2), j = clampi(j0 - 2 + tj, -1, g.Ny + 2);
        va[k] = xu[off(g, fx(g, i), cy(g, j))];
        vb[k] = xv[off(g, cx(g, i), fy(g, j))];
    }
}
template <typename T>
__device__ __forceinline__ void store_gtile(T (*tu)[GTW], T (*tv)[GTW], const T (&va)[GT_ITER], const T (&vb)[GT_ITER]) {
    const int tid = threadIdx.y * BX + threadIdx.x;
#pragma unroll
    for (int k = 0; k < GT_ITER; k++) {   // idle lanes write the last element's value into it again (no branch)
        const int e0 = tid + k * NT, e = e0 < GTH * GTW ? e0 : GTH * GTW - 1;
        const int tj = e / GTW, ti = e - tj * GTW;
        tu[tj][ti] = va[k];
        tv[tj][ti] = vb[k];
    }
}
template <typename T, bool PRO = false>
__device__ __forceinline__ void load_gtile(const KG &g, const T *__restrict__ xu, const T *__restrict__ xv, T (*tu)[GTW], T (*tv)[GTW],
                                           int i0, int j0, const ProSrc<T> *pr = nullptr) {
    const int tid = threadIdx.y * BX + threadIdx.x;
    for (int e = tid; e < GTH * GTW; e += NT) {
        const int tj = e / GTW, ti = e - tj * GTW;
        // every column a real point reads lies in [−1, N + 2]; beyond it the clamp only keeps idle elements in range
        const int i = clampi(i0 - 2 + ti, -1, g.Nx + 2), j = clampi(j0 - 2 + tj, -1, g.Ny + 2);
        const int ui = fx(g, i), uj = cy(g, j), vi = cx(g, i), vj = fy(g, j);
        T a = xu[off(g, ui, uj)], b = xv[off(g, vi, vj)];
        if (PRO) {   // k_prolongate's update of the points 1..Nx × 1..Ny, with its operands and order
            if (ui >= 1 && ui <= g.Nx && uj >= 1 && uj <= g.Ny) {
                const T p = prolong_u<T>(g, pr->cg, pr->hx, pr->hy, pr->cact, pr->cu, ui, uj);
                a += (pr->fact[off(g, ui, uj)] & 1) ? p : T(0);
            }
            if (vi >= 1 && vi <= g.Nx && vj >= 1 && vj <= g.Ny) {
                const T p = prolong_v<T>(g, pr->cg, pr->hx, pr->hy, pr->cact, pr->cv, vi, vj);
                b += (pr->fact[off(g, vi, vj)] & 2) ? p : T(0);
            }
        }
        tu[tj][ti] = a;
        tv[tj][ti] = b;
    }
}
template <typename T, int TX, int TY, bool RES, bool PRO = false>
__device__ __forceinline__ void gp_pass_t(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                          const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                          T *__restrict__ ov, const ProSrc<T> *pr = nullptr) {
    __shared__ T tu[GTH][GTW], tv[GTH][GTW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * BY + 1;
    load_gtile<T, PRO>(g, xu, xv, tu, tv, i0, j0, pr);
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i > g.Nx || j > g.Ny) return;
    auto XU = [&](int dx, int dy) { return tu[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    auto XV = [&](int dx, int dy) { return tv[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    const long o = off(g, i, j), P = gplane(g);
    const T qu = gp_row_d<T, 0>(g, pk, i, j, XU, XV);
    const T qv = gp_row_d<T, 1>(g, pk, i, j, XU, XV);
    if (RES) {
        ou[o] = bu[o] - qu;
        ov[o] = bv[o] - qv;
    } else {
        const T su = st[GS_PLANES * P + o], sv = st[(GS_PLANES + 1) * P + o];
        ou[o] = XU(0, 0) + (bu[o] - qu) * su;
        ov[o] = XV(0, 0) + (bv[o] - qv) * sv;
    }
}
// gp_pass_t for the latency-bound levels (a few thousand to tens of thousands of points, where a launch is a handful
// of round trips to memory): every tile element's loads and the row's couplings are issued together and waited for
// once before the barrier. Same values and order, so the results are gp_pass_t's bit for bit. On the bandwidth-bound
// levels the 62 couplings held in registers cost occupancy, so those keep gp_pass_t.
template <typename T, int TX, int TY, bool RES>
__device__ __forceinline__ void gp_pass_tl(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                           const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ou,
                                           T *__restrict__ ov) {
    __shared__ T tu[GTH][GTW], tv[GTH][GTW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * BY + 1;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool live = i <= g.Nx && j <= g.Ny;
    const long o = live ? off(g, i, j) : 0, P = gplane(g);
    T va[GT_ITER], vb[GT_ITER];
    fetch_gtile<T>(g, xu, xv, i0, j0, va, vb);
    // the row's couplings, right-hand side and 1/ℓ1 while the tile's loads are in flight (branch-free: a thread past
    // the level's edge loads point (1, 1)'s and discards them, so the wait before the tile is written counts only the
    // tile's loads)
    T cu[31], cv[31];
    const int ic = live ? i : 1, jc = live ? j : 1;
    const long oc = off(g, ic, jc);
    gp_coefs<T, 0>(g, pk, ic, jc, cu);
    gp_coefs<T, 1>(g, pk, ic, jc, cv);
    const T b_u = bu[oc], b_v = bv[oc];
    const T s_u = RES ? T(0) : st[GS_PLANES * P + oc], s_v = RES ? T(0) : st[(GS_PLANES + 1) * P + oc];
    store_gtile<T>(tu, tv, va, vb);
    keep_loaded(cu); keep_loaded(cv); keep_loaded(b_u); keep_loaded(b_v); keep_loaded(s_u); keep_loaded(s_v);
    __syncthreads();
    if (!live) return;
    auto XU = [&](int dx, int dy) { return tu[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    auto XV = [&](int dx, int dy) { return tv[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    const T qu = gp_row_cf<T, 0>(cu, XU, XV);
    const T qv = gp_row_cf<T, 1>(cv, XU, XV);
    if (RES) {
        ou[o] = b_u - qu;
        ov[o] = b_v - qv;
    } else {
        ou[o] = XU(0, 0) + (b_u - qu) * s_u;
        ov[o] = XV(0, 0) + (b_v - qv) * s_v;
    }
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_t(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                                   const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv,
                                                   T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    gp_pass_t<T, TX, TY, false>(g, st, pk, xu, xv, bu, bv, ou, ov);
}
// The first post-smoothing sweep with the prolongation folded into the tile load: every tile element k_prolongate
// would update gets x + P e with k_prolongate's arithmetic (the same helpers, no contraction), so the result is
// k_prolongate followed by k_gp_sweep_t, bit for bit, without the corrected x being stored and loaded again. The
// correction is formed per tile element, about twice per point; the coarse e and flags come from L2.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_pt(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                                    const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv,
                                                    T *__restrict__ ou, T *__restrict__ ov, ProSrc<T> pr, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    pr.cg.tx = TX; pr.cg.ty = TY;
    if (solve_done(gate)) return;
    gp_pass_t<T, TX, TY, false, true>(g, st, pk, xu, xv, bu, bv, ou, ov, &pr);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_residual_t(KG g, const T *__restrict__ pk, const T *__restrict__ xu, const T *__restrict__ xv,
                                                      const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ru,
                                                      T *__restrict__ rv, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    gp_pass_t<T, TX, TY, true>(g, nullptr, pk, xu, xv, bu, bv, ru, rv);
}

// k_gp_sweep_pre with x₁ = b·(1/ℓ1) formed once per element of a 68 × 8 LDS tile instead of at every stencil read (two
// loads and a product each); the same products, so the result is k_gp_sweep_pre's bit for bit.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_pre_t(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ bu,
                                                       const T *__restrict__ bv, T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    __shared__ T tu[GTH][GTW], tv[GTH][GTW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * BY + 1;
    const long P = gplane(g);
    const T *su_p = st + GS_PLANES * P, *sv_p = st + (GS_PLANES + 1) * P;
    const int tid = threadIdx.y * BX + threadIdx.x;
    for (int e = tid; e < GTH * GTW; e += NT) {
        const int tj = e / GTW, ti = e - tj * GTW;
        const int i = clampi(i0 - 2 + ti, -1, g.Nx + 2), j = clampi(j0 - 2 + tj, -1, g.Ny + 2);
        const long qu = off(g, fx(g, i), cy(g, j)), qv = off(g, cx(g, i), fy(g, j));
        tu[tj][ti] = bu[qu] * su_p[qu];
        tv[tj][ti] = bv[qv] * sv_p[qv];
    }
    __syncthreads();
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i > g.Nx || j > g.Ny) return;
    auto XU = [&](int dx, int dy) { return tu[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    auto XV = [&](int dx, int dy) { return tv[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    const long o = off(g, i, j);
    const T su = su_p[o], sv = sv_p[o];
    const T qu = gp_row_d<T, 0>(g, pk, i, j, XU, XV);
    const T qv = gp_row_d<T, 1>(g, pk, i, j, XU, XV);
    ou[o] = bu[o] * su + (bu[o] - qu) * su;
    ov[o] = bv[o] * sv + (bv[o] - qv) * sv;
}

// k_gp_sweep_t and k_gp_residual_t for the latency-bound levels (gp_pass_tl): results bit for bit the same.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_tl(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ xu,
                                                    const T *__restrict__ xv, const T *__restrict__ bu, const T *__restrict__ bv,
                                                    T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    gp_pass_tl<T, TX, TY, false>(g, st, pk, xu, xv, bu, bv, ou, ov);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_residual_tl(KG g, const T *__restrict__ pk, const T *__restrict__ xu, const T *__restrict__ xv,
                                                       const T *__restrict__ bu, const T *__restrict__ bv, T *__restrict__ ru,
                                                       T *__restrict__ rv, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    gp_pass_tl<T, TX, TY, true>(g, nullptr, pk, xu, xv, bu, bv, ru, rv);
}

// k_gp_sweep_pre_t for the latency-bound levels, as gp_pass_tl: the same values and order, bit for bit.
template <typename T, int TX, int TY>
__global__ __launch_bounds__(NT) void k_gp_sweep_pre_tl(KG g, const T *__restrict__ st, const T *__restrict__ pk, const T *__restrict__ bu,
                                                        const T *__restrict__ bv, T *__restrict__ ou, T *__restrict__ ov, const Scal *gate) {
    g.tx = TX; g.ty = TY;
    if (solve_done(gate)) return;
    __shared__ T tu[GTH][GTW], tv[GTH][GTW];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int i0 = bx_ * BX + 1, j0 = by_ * BY + 1;
    const long P = gplane(g);
    const T *su_p = st + GS_PLANES * P, *sv_p = st + (GS_PLANES + 1) * P;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool live = i <= g.Nx && j <= g.Ny;
    const long o = live ? off(g, i, j) : 0;
    // b and 1/ℓ1 of the tile elements (all loads of a thread issued together), then the row's couplings, right-hand
    // side and 1/ℓ1 while those are in flight; x₁ = b·(1/ℓ1) is formed as the tile is written
    const int tid = threadIdx.y * BX + threadIdx.x;
    T tb_u[GT_ITER], ts_u[GT_ITER], tb_v[GT_ITER], ts_v[GT_ITER];
#pragma unroll
    for (int k = 0; k < GT_ITER; k++) {
        const int e0 = tid + k * NT, e = e0 < GTH * GTW ? e0 : GTH * GTW - 1;   // idle lanes reread the last element
        const int tj = e / GTW, ti = e - tj * GTW;
        const int ii = clampi(i0 - 2 + ti, -1, g.Nx + 2), jj = clampi(j0 - 2 + tj, -1, g.Ny + 2);
        const long qu = off(g, fx(g, ii), cy(g, jj)), qv = off(g, cx(g, ii), fy(g, jj));
        tb_u[k] = bu[qu]; ts_u[k] = su_p[qu];
        tb_v[k] = bv[qv]; ts_v[k] = sv_p[qv];
    }
    // (branch-free: a thread past the level's edge loads point (1, 1)'s and discards them)
    T cu[31], cv[31];
    const int ic = live ? i : 1, jc = live ? j : 1;
    const long oc = off(g, ic, jc);
    gp_coefs<T, 0>(g, pk, ic, jc, cu);
    gp_coefs<T, 1>(g, pk, ic, jc, cv);
    const T b_u = bu[oc], b_v = bv[oc], s_u = su_p[oc], s_v = sv_p[oc];
#pragma unroll
    for (int k = 0; k < GT_ITER; k++) {   // idle lanes write the last element's value into it again (no branch)
        const int e0 = tid + k * NT, e = e0 < GTH * GTW ? e0 : GTH * GTW - 1;
        const int tj = e / GTW, ti = e - tj * GTW;
        tu[tj][ti] = tb_u[k] * ts_u[k];
        tv[tj][ti] = tb_v[k] * ts_v[k];
    }
    keep_loaded(cu); keep_loaded(cv); keep_loaded(b_u); keep_loaded(b_v); keep_loaded(s_u); keep_loaded(s_v);
    __syncthreads();
    if (!live) return;
    auto XU = [&](int dx, int dy) { return tu[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    auto XV = [&](int dx, int dy) { return tv[j - j0 + 2 + dy][i - i0 + 2 + dx]; };
    const T qu = gp_row_cf<T, 0>(cu, XU, XV);
    const T qv = gp_row_cf<T, 1>(cv, XU, XV);
    ou[o] = b_u * s_u + (b_u - qu) * s_u;
    ov[o] = b_v * s_v + (b_v - qv) * s_v;
}

// The residual of a latency-bound packed Galerkin level and its restriction to the level below in one launch. A block
// owns an RRX × RRY tile of the coarse level. Each of its threads forms one fine residual r = b − A x that the tile's
// restriction reads — at the fine index k_restrict reads it (r_index's map) and masked with the fine activity as
// k_restrict masks it — into LDS, with every load it needs (the row's 31 couplings, the 31 iterate values, b and the
// activity byte) issued at once; then 2·RRX·RRY threads form b = R r of the tile's points with k_restrict's weights,
// order and coarse activity and store it. The residual arithmetic is k_gp_residual's (gp_row_cf in gp_row_d's order)
// and the restriction restrict_u / restrict_v's, so b is k_gp_residual_tl + k_restrict's bit for bit, with one
// launch and one round trip to memory fewer; r itself is not stored (only the restriction reads it).
constexpr int RRX = 32, RRY = 2;                          // coarse tile of a block
constexpr int RRU_W = 2 * RRX + 1, RRU_H = 2 * RRY + 2;   // fine u residuals of the tile (both directions halved)
constexpr int RRV_W = 2 * RRX + 2, RRV_H = 2 * RRY + 1;   // fine v residuals
constexpr int RR_THREADS = 768;                           // ≥ RRU_W·RRU_H + RRV_W·RRV_H = 720 slots, one per thread
static_assert(RRU_W * RRU_H + RRV_W * RRV_H <= RR_THREADS, "one residual slot per thread");
// r_index's fine index for an unmapped one, x = 2I − 1 + a on a halved direction (and I itself on an un-halved one)
__device__ __forceinline__ int rr_map(int x, int N, int bounded, int face, int h) {
    if (!h) return x;
    if (!bounded) return wrapi(x, N);
    if (face) return x > 1 ? x : 1;
    return clampi(x, 1, N);
}
template <typename T, int TX, int TY>
__global__ __launch_bounds__(RR_THREADS) void k_gp_residual_restrict(KG fg, const T *__restrict__ pk, const T *__restrict__ xu,
                                                                    const T *__restrict__ xv, const T *__restrict__ bu,
                                                                    const T *__restrict__ bv, const uint8_t *__restrict__ fact,
                                                                    KG cg, int hx, int hy, const uint8_t *__restrict__ cact,
                                                                    T *__restrict__ cbu, T *__restrict__ cbv, const Scal *gate) {
    fg.tx = TX; fg.ty = TY;
    cg.tx = TX; cg.ty = TY;
    if (solve_done(gate)) return;
    __shared__ T su[RRU_H][RRU_W], sv[RRV_H][RRV_W];
    int bx_, by_;
    xcd_tile(bx_, by_);
    const int I0 = bx_ * RRX + 1, J0 = by_ * RRY + 1;
    // unmapped fine index of slot 0 along each direction, and the slot extents of the two components
    const int xo = hx ? 2 * I0 - 2 : I0, yo = hy ? 2 * J0 - 2 : J0;
    const int uw = hx ? RRU_W : RRX, uh = hy ? RRU_H : RRY, vw = hx ? RRV_W : RRX, vh = hy ? RRV_H : RRY;
    const int tid = threadIdx.x, nu = uw * uh;
    const bool isu = tid < nu, live = tid < nu + vw * vh;
    const int t = isu ? tid : tid - nu, w = isu ? uw : vw;
    const int si = live ? t % w : 0, sj = live ? t / w : 0;
    // unmapped indices beyond the ones a coarse point of the level reads are clamped, so every map and address stays
    // in range (a packed level has at least 16 points a side, so the clamp never moves a slot that is read)
    const int xi = clampi(xo + si, -fg.Nx, 2 * fg.Nx), yj = clampi(yo + sj, -fg.Ny, 2 * fg.Ny);
    const int mi = rr_map(xi, fg.Nx, fg.tx, isu ? 1 : 0, hx), mj = rr_map(yj, fg.Ny, fg.ty, isu ? 0 : 1, hy);
    // the residual is formed at an in-range point (the k_gp_residual range); an out-of-range one is a wall, inactive
    const int ci = clampi(mi, 1, fg.Nx), cj = clampi(mj, 1, fg.Ny);
    const long o = off(fg, ci, cj);
    const bool act = isu ? uact(fg, fact, mi, mj) : vact(fg, fact, mi, mj);
    T cf[31];
    if (isu) gp_coefs<T, 0>(fg, pk, ci, cj, cf);
    else gp_coefs<T, 1>(fg, pk, ci, cj, cf);
    const T b0 = isu ? bu[o] : bv[o];
    auto XU = [&](int dx, int dy) { return xu[off(fg, fx(fg, ci + dx), cy(fg, cj + dy))]; };
    auto XV = [&](int dx, int dy) { return xv[off(fg, cx(fg, ci + dx), fy(fg, cj + dy))]; };
    const T q = isu ? gp_row_cf<T, 0>(cf, XU, XV) : gp_row_cf<T, 1>(cf, XU, XV);
    const T r = b0 - q;
    if (live) {
        if (isu) su[sj][si] = act ? r : T(0);
        else sv[sj][si] = act ? r : T(0);
    }
    __syncthreads();
    if (tid >= 2 * RRX * RRY) return;
    const int comp = tid / (RRX * RRY), pp = tid % (RRX * RRY);
    const int I = I0 + pp % RRX, J = J0 + pp / RRX;
    if (I > cg.Nx || J > cg.Ny) return;
    const long oc = off(cg, I, J);
    T sum = T(0);
    if (comp == 0) {   // restrict_u's loop
#pragma unroll
        for (int a = -1; a <= 1; a++)
#pragma unroll
            for (int b = -1; b <= 2; b++) {
                const bool on = (hx || a == 0) && (hy || b == 0);
                const T wgt = r_weight<T>(I, a, cg.Nx, fg.tx, 1, hx) * r_weight<T>(J, b, cg.Ny, fg.ty, 0, hy);
                const T val = su[(hy ? 2 * J - 1 + b : J) - yo][(hx ? 2 * I - 1 + a : I) - xo];
                if (on) sum += wgt * val;
            }
        cbu[oc] = (cact[oc] & 1) ? sum : T(0);
    } else {           // restrict_v's loop
#pragma unroll
        for (int a = -1; a <= 2; a++)
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                const bool on = (hx || a == 0) && (hy || b == 0);
                const T wgt = r_weight<T>(I, a, cg.Nx, fg.tx, 0, hx) * r_weight<T>(J, b, cg.Ny, fg.ty, 1, hy);
                const T val = sv[(hy ? 2 * J - 1 + b : J) - yo][(hx ? 2 * I - 1 + a : I) - xo];
                if (on) sum += wgt * val;
            }
        cbv[oc] = (cact[oc] & 2) ? sum : T(0);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Coarsest level: the Galerkin operator as a dense matrix, inverted once per build (Gauss-Jordan without pivoting:
// the operator is symmetric positive semidefinite, an inactive row is an identity row). A pivot that falls below
// a relative tolerance of its original diagonal marks a null direction (a floating, doubly periodic domain has the
// rigid translations): that unknown is left out and its row and column of the result are zero, which leaves the
// inverse of the remaining block — a symmetric positive semidefinite generalized inverse. Unknown k < n/2 is u at
// (k % Nx + 1, k / Nx + 1), k ≥ n/2 is v likewise; n = 2·Nx·Ny.
// ---------------------------------------------------------------------------------------------------------------
constexpr int GDENSE_MAX = 512;   // largest coarsest system inverted densely; above it the coarsest level is swept
constexpr int GDENSE_STOP = 128;  // a Galerkin hierarchy ends at the first coarse level with at most this many unknowns

__device__ __forceinline__ int g_unknown(const KG &g, int comp, int i, int j) { return comp * g.Nx * g.Ny + (j - 1) * g.Nx + (i - 1); }

// The inverse is formed and stored in FP64 whatever the V-cycle's precision (the mixed form keeps the coarsest-grid
// solve in FP64); a null direction is recognised relative to the rounding of the stencil it was assembled from.
// Systems of up to GDENSE_LDS unknowns are eliminated in LDS (n² doubles of dynamic shared memory) and only the result
// is written to M; larger ones are eliminated in M itself.
constexpr int GDENSE_LDS = 128;
// Gauss-Jordan of an n × n system (n ≤ GDENSE_LDS) held in registers: with GINV_ROWS × GDENSE_LDS threads, thread t keeps
// column j = t mod GDENSE_LDS of rows t / GDENSE_LDS + GINV_ROWS·r (a wave's 64 lanes share one row group, so row
// indices are wave-uniform). Step k reads row k, column k and 1/M[k][k] as step k − 1 left them from LDS (double
// buffered, one barrier per step): the owners of row k + 1 and column k + 1 publish them, and 1/M[k+1][k+1], as the
// step produces them. Every element is updated with the operations, in the order, of the shared-memory elimination
// below — row k: (j == k ? 1 : M[k][j])·(1/p); other rows: M[i][j] − M[i][k]·m_kj; column k: (−0) − M[i][k]·(1/p),
// which is −(M[i][k]·(1/p)) bit for bit, zeros' signs included — so the inverse is that loop's bit for bit.
// The update of the other rows has no selects (the pivot row is overwritten afterwards by its owners alone): the
// FP64 multiply and subtract of 16 rows per thread per step set the pace.
constexpr int GINV_ROWS = TAIL_THREADS / GDENSE_LDS;
constexpr int GINV_R = GDENSE_LDS / GINV_ROWS;   // rows per thread
__device__ __forceinline__ void g_inverse_regs(const double *__restrict__ Ml, double *__restrict__ Mout, int n, double tol,
                                               const double *dorig, unsigned char *sing) {
    __shared__ double rowb[2][GDENSE_LDS], colb[2][GDENSE_LDS], pivb[2];
    const int tid = threadIdx.x, j = tid % GDENSE_LDS;
    const int rb = __builtin_amdgcn_readfirstlane(tid / GDENSE_LDS);
    const bool cj = j < n;
    double R[GINV_R];
#pragma unroll
    for (int r = 0; r < GINV_R; r++) {
        const int i = rb + GINV_ROWS * r;
        R[r] = (cj && i < n) ? Ml[(long)i * n + j] : 0.0;
    }
    if (rb == 0 && cj) rowb[0][j] = R[0];
    if (tid == 0) pivb[0] = 1.0 / R[0];
    if (j == 0) {
#pragma unroll
        for (int r = 0; r < GINV_R; r++) {
            const int i = rb + GINV_ROWS * r;
            if (i < n) colb[0][i] = R[r];
        }
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        const int b = k & 1, nb = b ^ 1, k1 = k + 1;
        // every read of the step in flight together (column entries at i ≥ n and row entries at j ≥ n are never used)
        const double pk = rowb[b][k], piv = pivb[b], rkj = rowb[b][j];
        double f[GINV_R];
#pragma unroll
        for (int r = 0; r < GINV_R; r++) f[r] = colb[b][rb + GINV_ROWS * r];
        if (!(pk > tol * fabs(dorig[k]))) {   // a null direction: unknown k is left out
            if (tid == 0) sing[k] = 1;
        } else {
            const double mkj = (j == k ? 1.0 : rkj) * piv;
            if (j == k) {
#pragma unroll
                for (int r = 0; r < GINV_R; r++)
                    if (rb + GINV_ROWS * r != k) R[r] = -0.0;
            }
#pragma unroll
            for (int r = 0; r < GINV_R; r++) R[r] = R[r] - f[r] * mkj;   // rows ≥ n are padding, never read
            if (rb == k % GINV_ROWS) {   // the pivot row: (j == k ? 1 : M[k][j])·(1/p)
                const int kr = k / GINV_ROWS;
#pragma unroll
                for (int r = 0; r < GINV_R; r++) R[r] = r == kr ? mkj : R[r];
            }
        }
        if (k1 < n) {   // row k + 1, its reciprocal pivot and column k + 1 as this step leaves them
            if (rb == k1 % GINV_ROWS) {
                const int kr = k1 / GINV_ROWS;
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < GINV_R; r++) v = r == kr ? R[r] : v;
                if (cj) rowb[nb][j] = v;
                if (j == k1) pivb[nb] = 1.0 / v;   // unused when step k + 1 leaves the unknown out
            }
            if (j == k1) {
#pragma unroll
                for (int r = 0; r < GINV_R; r++) {
                    const int i = rb + GINV_ROWS * r;
                    if (i < n) colb[nb][i] = R[r];
                }
            }
        }
        __syncthreads();
    }
    // null directions dropped; stored column-major (element (r, c) at c·n + r) for coalesced solves
    if (cj) {
#pragma unroll
        for (int r = 0; r < GINV_R; r++) {
            const int i = rb + GINV_ROWS * r;
            if (i < n) Mout[(long)j * n + i] = (sing[i] || sing[j]) ? 0.0 : R[r];
        }
    }
}

// Row RC at (i, j) of the dense coarsest operator: its 31 couplings added into their columns of Mrow in pattern order
// (an out-of-range column carries a zero coupling; on a small Periodic level two offsets may wrap onto one column and
// add in that order). The couplings are loaded first, all in flight together, then added.
template <typename T, typename TS, int RC>
__device__ __forceinline__ void g_inverse_row(const KG &g, const TS *__restrict__ st, long P, int i, int j, T *Mrow) {
    constexpr int base = gblk(RC, 0).base;
    const long o = off(g, i, j);
    T cf[31];
#pragma unroll
    for (int k = 0; k < 31; k++) cf[k] = (T)st[(base + k) * P + o];
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
        constexpr GBlk B0 = gblk(RC, 0), B1 = gblk(RC, 1);
        const GBlk b = cc == 0 ? B0 : B1;
#pragma unroll
        for (int dx = b.x0; dx <= b.x1; dx++)
#pragma unroll
            for (int dy = b.y0; dy <= b.y1; dy++) {
                const T a = cf[gidx(b, dx, dy) - base];
                if (a == T(0)) continue;
                int ic = i + dx, jc = j + dy;
                if (!g.tx) ic = wrapi(ic, g.Nx);
                if (!g.ty) jc = wrapi(jc, g.Ny);
                if (ic < 1 || ic > g.Nx || jc < 1 || jc > g.Ny) continue;
                Mrow[g_unknown(g, cc, ic, jc)] += a;
            }
    }
}

template <typename TS, int TX, int TY>
__global__ __launch_bounds__(TAIL_THREADS) void k_g_inverse(KG g, const TS *__restrict__ st, double *__restrict__ Mout, int regs) {
    using T = double;
    g.tx = TX; g.ty = TY;
    extern __shared__ __align__(16) double gsm[];
    __shared__ T fcol[GDENSE_MAX];
    __shared__ T dorig[GDENSE_MAX];
    __shared__ unsigned char sing[GDENSE_MAX];
    __shared__ T piv;
    __shared__ int skip;
    const T tol = sizeof(TS) == 4 ? T(1e-5) : T(1e-10);
    const int n = 2 * g.Nx * g.Ny, tid = threadIdx.x, nth = blockDim.x;
    const bool lds = n <= GDENSE_LDS;
    T *M = lds ? gsm : Mout;
    const long P = gplane(g);
    for (int e = tid; e < n * n; e += nth) M[e] = T(0);
    __syncthreads();
    // assemble: row r of component rc at (i, j) from its 31 couplings (out-of-range columns carry zero couplings)
    for (int r = tid; r < n; r += nth) {
        const int rc = r >= n / 2, k = r - rc * (n / 2), i = k % g.Nx + 1, j = k / g.Nx + 1;
        if (rc == 0) g_inverse_row<T, TS, 0>(g, st, P, i, j, M + (long)r * n);
        else g_inverse_row<T, TS, 1>(g, st, P, i, j, M + (long)r * n);
    }
    __syncthreads();
    for (int i = tid; i < n; i += nth) {
        dorig[i] = M[(long)i * n + i];
        sing[i] = 0;
    }
    __syncthreads();
    if (regs && lds && nth == GINV_ROWS * GDENSE_LDS) {
        g_inverse_regs(M, Mout, n, tol, dorig, sing);
        return;
    }
    for (int k = 0; k < n; k++) {
        if (tid == 0) {
            const T p = M[(long)k * n + k];
            skip = !(p > tol * fabs(dorig[k]));
            if (skip) {
                sing[k] = 1;
            } else {
                piv = T(1) / p;
                M[(long)k * n + k] = T(1);
            }
        }
        __syncthreads();
        const int sk = skip;   // every thread reads the flag before thread 0 may write the next one
        __syncthreads();
        if (sk) continue;
        for (int j = tid; j < n; j += nth) M[(long)k * n + j] *= piv;
        for (int i = tid; i < n; i += nth) fcol[i] = M[(long)i * n + k];
        __syncthreads();
        for (int e = tid; e < n * n; e += nth) {
            const int i = e / n, j = e % n;
            if (i == k) continue;
            const T f = fcol[i];
            M[e] = (j == k) ? -f * M[(long)k * n + k] : M[e] - f * M[(long)k * n + j];
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += nth)
        if (sing[e / n] || sing[e % n]) M[e] = T(0);
    __syncthreads();
    // store column-major (element (r, c) at c·n + r) so that the solves read it with coalesced loads
    if (lds) {
        for (int e = tid; e < n * n; e += nth) {
            const int c = e / n, r = e % n;
            Mout[e] = M[(long)r * n + c];
        }
        return;
    }
    for (int e = tid; e < n * n; e += nth) {
        const int i = e / n, j = e % n;
        if (i < j) {
            const T a = M[(long)i * n + j];
            M[(long)i * n + j] = M[(long)j * n + i];
            M[(long)j * n + i] = a;
        }
    }
}

// x = M⁻¹ b on the coarsest level (one workgroup), in FP64 (b promoted exactly, x rounded once to the level's precision)
template <typename T>
__global__ __launch_bounds__(TAIL_THREADS) void k_g_dense_solve(KG g, const double *__restrict__ M, const T *__restrict__ bu,
                                                                const T *__restrict__ bv, T *__restrict__ xu, T *__restrict__ xv,
                                                                const Scal *gate) {
    if (solve_done(gate)) return;
    const int n = 2 * g.Nx * g.Ny;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double acc = 0.0;
        for (int c = 0; c < n; c++) {
            const int cc = c >= n / 2, k = c - cc * (n / 2);
            const long o = off(g, k % g.Nx + 1, k / g.Nx + 1);
            acc += M[(long)c * n + r] * (double)(cc ? bv[o] : bu[o]);
        }
        const int rc = r >= n / 2, k = r - rc * (n / 2);
        const long o = off(g, k % g.Nx + 1, k / g.Nx + 1);
        (rc ? xv : xu)[o] = (T)acc;
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Coarse tail with Galerkin operators: the levels below tail_start of a V-cycle in one 1024-thread workgroup.
// Vectors live in LDS (two x buffers for Jacobi ping-pong, b, r: 8 values and 1 flag byte per point); the stencil
// planes of the smallest levels are copied into LDS too (o_st ≥ 0), the others are read from global memory
// (L2-resident at these sizes). The coarsest level is solved with its dense inverse when it has one.
// ---------------------------------------------------------------------------------------------------------------
// Compact copy of a level's stencil for the LDS tail: the GS_ALLOC planes with pitch Nx + 1 and no padding rows,
// plane k of point (i, j) at k·(Nx + 1)(Ny + 1) + (j − 1)(Nx + 1) + i − 1 — the layout the tail keeps in LDS, so
// loading it is one contiguous 16-byte copy, and a level read from global memory is read with coalesced loads.
template <typename T>
__global__ __launch_bounds__(NT) void k_g_compact(KG g, const T *__restrict__ st, T *__restrict__ cst) {
    NK_IJ
    if (i > g.Nx + 1 || j > g.Ny + 1) return;
    const long P = gplane(g), o = off(g, i, j);
    const long NP = (long)(g.Nx + 1) * (g.Ny + 1), q = (long)(j - 1) * (g.Nx + 1) + (i - 1);
    const bool row = i <= g.Nx && j <= g.Ny;
#pragma unroll 8
    for (int k = 0; k < GS_ALLOC; k++) cst[k * NP + q] = row ? st[k * P + o] : T(0);
}

template <typename T> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int n = 4; };
template <> struct V16<double> { using type = double2; static constexpr int n = 2; };

constexpr int GTAIL_MAX = 12;

template <typename T> struct GTailLevel {
    KG lg;                  // LDS geometry (pitch = Nx + 1)
    int gpitch;             // pitch of the level's global planes
    int hx, hy;             // halved relative to the next finer level
    const T *cst;           // compact stencil (k_g_compact layout, global)
    const uint8_t *act;     // activity flags (global)
    int o_x0u, o_x0v, o_x1u, o_x1v, o_bu, o_bv, o_ru, o_rv, o_act;   // LDS offsets (elements of T; bytes for o_act)
    int o_st;               // LDS offset of a copy of the compact stencil (16-byte aligned), or −1 (read cst)
};

template <typename T> struct GTailDesc {
    int n, sweeps, coarsest;
    int dgroups, o_dp;      // dense solve: column groups and the LDS byte offset of their FP64 partial products (G·n),
                            // followed by the coarsest right-hand side in FP64 (n, unknown order)
    const T *gbu, *gbv;     // right-hand side of the top tail level (global)
    T *gxu, *gxv;           // result of the top tail level (global)
    const double *cinv;     // FP64 dense inverse of the coarsest level, or null (sweeps)
    GTailLevel<T> lv[GTAIL_MAX];
};

// Development aid: built with -DNK_TAIL_TIMING, thread 0 of the tail stamps the device clock (100 MHz) after the
// load and after every step of the cycle and prints the 41st call's step times. Off by default.
#ifdef NK_TAIL_TIMING
__device__ unsigned int nk_tail_calls = 0;
#define NK_TT(k) do { const int k_ = (k); if (threadIdx.x == 0 && k_ < 64) tt[k_] = wall_clock64(); } while (0)
#else
#define NK_TT(k) do { } while (0)
#endif

// The cycle is one loop over its steps (sweeps, residual, restriction, coarsest solve, prolongation), so each
// operation's code appears once, and the rows of a pass are spread one per thread (u rows, then v rows), so a 16²
// level keeps 512 threads busy instead of 256. Per row the arithmetic is that of g_row, k_restrict and k_prolongate.

enum TailOp : int { TAIL_SWEEP = 0, TAIL_RESIDUAL = 1, TAIL_RESTRICT = 2, TAIL_DENSE = 3, TAIL_PROLONG = 4 };

template <typename T, int TX, int TY>
__global__ __launch_bounds__(TAIL_THREADS) void k_g_tail(GTailDesc<T> d, const Scal *gate) {
    if (solve_done(gate)) return;
#ifdef NK_TAIL_TIMING
    unsigned long long tt[64];
    int nt = 0;
#endif
    NK_TT(nt++);
    auto geom = [](const KG &k) { KG g = k; g.tx = TX; g.ty = TY; return g; };
    extern __shared__ __align__(16) unsigned char smem[];
    T *sm = reinterpret_cast<T *>(smem);
    const int tid = threadIdx.x, nth = blockDim.x;
    // the coarsest right-hand side in FP64 in unknown order, written by the restriction into the coarsest level, so
    // the dense solve reads b[c] directly (the promotion to FP64 is exact)
    const int ncoarse = 2 * d.lv[d.n - 1].lg.Nx * d.lv[d.n - 1].lg.Ny;
    double *bvec = d.cinv ? reinterpret_cast<double *>(smem + d.o_dp) + d.dgroups * ncoarse : nullptr;
    for (int l = 0; l < d.n; l++) {
        const GTailLevel<T> &L = d.lv[l];
        const int nx = L.lg.Nx + 1, ny = L.lg.Ny + 1;
        for (int p = tid; p < nx * ny; p += nth) {
            const int i = p % nx, j = p / nx;
            const long go = (long)j * L.gpitch + i;
            const int lo = j * L.lg.pitch + i;
            smem[L.o_act + lo] = L.act[go];
            sm[L.o_x0u + lo] = T(0); sm[L.o_x0v + lo] = T(0);
            sm[L.o_x1u + lo] = T(0); sm[L.o_x1v + lo] = T(0);
            sm[L.o_ru + lo] = T(0); sm[L.o_rv + lo] = T(0);
            sm[L.o_bu + lo] = l == 0 ? d.gbu[go] : T(0);
            sm[L.o_bv + lo] = l == 0 ? d.gbv[go] : T(0);
        }
    }
    // compact stencils of the LDS-resident levels: contiguous 16-byte copies
    for (int l = 0; l < d.n; l++) {
        const GTailLevel<T> &L = d.lv[l];
        if (L.o_st < 0) continue;
        using V = typename V16<T>::type;
        const int nv = (L.lg.Nx + 1) * (L.lg.Ny + 1) * GS_ALLOC / V16<T>::n;
        const V *src = reinterpret_cast<const V *>(L.cst);
        V *dst = reinterpret_cast<V *>(sm + L.o_st);
#pragma unroll 4
        for (int q = tid; q < nv; q += nth) dst[q] = src[q];
    }
    __syncthreads();
    NK_TT(nt++);
    unsigned cur = 0;     // bit l: which x buffer holds level l's iterate (0: x0, 1: x1); a register, not an array
    auto C_ = [&](int l) { return (int)((cur >> l) & 1u); };
    auto X = [&](int l, int which, int comp) -> T * {
        const GTailLevel<T> &L = d.lv[l];
        return sm + (which == 0 ? (comp == 0 ? L.o_x0u : L.o_x0v) : (comp == 0 ? L.o_x1u : L.o_x1v));
    };
    // One pass over the level's rows with the compact stencil at stb (LDS copy or global): an ℓ1 sweep
    // o = x + (b − A x)/ℓ1 (mode 0: o = b/ℓ1, the sweep from zero; mode 1: the sweep), or the residual o = b − A x
    // (mode 2). Called with a pointer of one address space per call site, so the stencil loads are LDS or global
    // loads, not flat ones.
    auto rows = [&](const T *stb, const KG &g, const T *xu, const T *xv, const T *bu, const T *bv, T *ou, T *ov,
                    int mode) __attribute__((always_inline)) {
        const long NP = (long)(g.Nx + 1) * (long)(g.Ny + 1);
        const int npts = g.Nx * g.Ny;
        for (int r = tid; r < 2 * npts; r += nth) {
            const int rc = r >= npts, p = r - rc * npts;
            const int i = p % g.Nx + 1, j = p / g.Nx + 1;
            const long o = off(g, i, j);   // pitch Nx + 1: also the point's index in the compact stencil
            const T s = stb[(GS_PLANES + rc) * NP + o];
            const T b = rc ? bv[o] : bu[o];
            T out;
            if (mode == 0) {
                out = b * s;
            } else {
                const T q = rc ? g_row<T, 1>(g, stb, NP, o, i, j, xu, xv) : g_row<T, 0>(g, stb, NP, o, i, j, xu, xv);
                out = mode == 2 ? b - q : (rc ? xv[o] : xu[o]) + (b - q) * s;
            }
            (rc ? ov : ou)[o] = out;
        }
    };
    const int S = d.sweeps, nl = d.n;
    const int per_down = S + 2, n_down = (nl - 1) * per_down;
    const int n_coarse = d.cinv ? 1 : d.coarsest;
    const int per_up = S + 1, n_steps = n_down + n_coarse + (nl - 1) * per_up;
    for (int step = 0; step < n_steps; step++) {
        int l, op, z = 0;
        if (step < n_down) {            // down leg: S sweeps from zero, the residual, the restriction
            l = step / per_down;
            const int k = step - l * per_down;
            op = k < S ? TAIL_SWEEP : (k == S ? TAIL_RESIDUAL : TAIL_RESTRICT);
            z = k == 0;
        } else if (step < n_down + n_coarse) {   // coarsest level: dense inverse, or sweeps from zero
            l = nl - 1;
            op = d.cinv ? TAIL_DENSE : TAIL_SWEEP;
            z = step == n_down;
        } else {                        // up leg: the prolongation, S sweeps
            const int t = step - n_down - n_coarse, q = t / per_up;
            l = nl - 2 - q;
            op = t - q * per_up == 0 ? TAIL_PROLONG : TAIL_SWEEP;
        }
        const GTailLevel<T> &L = d.lv[l];
        const KG g = geom(L.lg);
        if (op == TAIL_SWEEP || op == TAIL_RESIDUAL) {
            const int src = C_(l), dst = op == TAIL_RESIDUAL ? src : (z ? src : 1 - src);
            T *ou = op == TAIL_RESIDUAL ? sm + L.o_ru : X(l, dst, 0), *ov = op == TAIL_RESIDUAL ? sm + L.o_rv : X(l, dst, 1);
            const int mode = op == TAIL_RESIDUAL ? 2 : (z ? 0 : 1);
            if (L.o_st >= 0) rows(sm + L.o_st, g, X(l, src, 0), X(l, src, 1), sm + L.o_bu, sm + L.o_bv, ou, ov, mode);
            else rows(L.cst, g, X(l, src, 0), X(l, src, 1), sm + L.o_bu, sm + L.o_bv, ou, ov, mode);
            if (op == TAIL_SWEEP) cur = (cur & ~(1u << l)) | ((unsigned)dst << l);
        } else if (op == TAIL_RESTRICT) {
            const GTailLevel<T> &C = d.lv[l + 1];
            const KG cg = geom(C.lg);
            const uint8_t *cact = smem + C.o_act, *fact = smem + L.o_act;
            const T *ru = sm + L.o_ru, *rv = sm + L.o_rv;
            const int cn = cg.Nx * cg.Ny;
            for (int r = tid; r < 2 * cn; r += nth) {
                const int rc = r >= cn, p = r - rc * cn;
                const int I = p % cg.Nx + 1, J = p / cg.Nx + 1;
                T sum = T(0);
                // fixed trip counts (as k_restrict): all reads issued together, off-support terms masked
                if (rc == 0) {   // u: Face along x, Center along y
#pragma unroll
                    for (int a = -1; a <= 1; a++)
#pragma unroll
                        for (int b = -1; b <= 2; b++) {
                            const bool on = (C.hx || a == 0) && (C.hy || b == 0);
                            const int i = r_index(I, a, g.Nx, g.tx, 1, C.hx), j = r_index(J, b, g.Ny, g.ty, 0, C.hy);
                            const T w = r_weight<T>(I, a, cg.Nx, g.tx, 1, C.hx) * r_weight<T>(J, b, cg.Ny, g.ty, 0, C.hy);
                            const T val = iu(g, fact, ru, i, j);
                            if (on) sum += w * val;
                        }
                } else {         // v: Center along x, Face along y
#pragma unroll
                    for (int a = -1; a <= 2; a++)
#pragma unroll
                        for (int b = -1; b <= 1; b++) {
                            const bool on = (C.hx || a == 0) && (C.hy || b == 0);
                            const int i = r_index(I, a, g.Nx, g.tx, 0, C.hx), j = r_index(J, b, g.Ny, g.ty, 1, C.hy);
                            const T w = r_weight<T>(I, a, cg.Nx, g.tx, 0, C.hx) * r_weight<T>(J, b, cg.Ny, g.ty, 1, C.hy);
                            const T val = iv(g, fact, rv, i, j);
                            if (on) sum += w * val;
                        }
                }
                const long o = off(cg, I, J);
                const T bval = (cact[o] & (rc ? 2 : 1)) ? sum : T(0);
                (rc ? sm + C.o_bv : sm + C.o_bu)[o] = bval;
                if (d.cinv && l + 1 == nl - 1) bvec[r] = (double)bval;   // row r is unknown r of the dense solve
            }
        } else if (op == TAIL_DENSE) {
            // x = M⁻¹ b: d.dgroups groups of n threads, group q summing columns [q·chunk, (q + 1)·chunk) of row r
            // (coalesced: M is column-major), then the groups' partials added in group order; in FP64 whatever the
            // V-cycle's precision (b promoted exactly, x rounded once to the level's precision)
            const int n = 2 * g.Nx * g.Ny, G = d.dgroups, chunk = (n + G - 1) / G;
            double *dp = reinterpret_cast<double *>(smem + d.o_dp);
            if (tid < G * n) {
                const int r = tid % n, q = tid / n;
                const int c0 = q * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
                double acc = 0.0;
                constexpr int U = 16;   // loads of a batch issued together
                for (int cb = c0; cb < c1; cb += U) {
                    double m[U];
#pragma unroll
                    for (int t = 0; t < U; t++) m[t] = cb + t < c1 ? d.cinv[(long)(cb + t) * n + r] : 0.0;
#pragma unroll
                    for (int t = 0; t < U; t++)
                        if (cb + t < c1) acc += m[t] * bvec[cb + t];
                }
                dp[q * n + r] = acc;
            }
            __syncthreads();
            T *xu = X(l, C_(l), 0), *xv = X(l, C_(l), 1);
            for (int r = tid; r < n; r += nth) {
                double acc = 0.0;
                for (int q = 0; q < G; q++) acc += dp[q * n + r];
                const int rc = r >= n / 2, k = r - rc * (n / 2);
                const long o = off(g, k % g.Nx + 1, k / g.Nx + 1);
                (rc ? xv : xu)[o] = (T)acc;
            }
        } else {   // TAIL_PROLONG: x at level l += P (x at level l + 1)
            const GTailLevel<T> &C = d.lv[l + 1];
            const KG cg = geom(C.lg);
            const uint8_t *cact = smem + C.o_act, *fact = smem + L.o_act;
            const T *cu = X(l + 1, C_(l + 1), 0), *cv = X(l + 1, C_(l + 1), 1);
            T *xu = X(l, C_(l), 0), *xv = X(l, C_(l), 1);
            const int npts = g.Nx * g.Ny;
            for (int r = tid; r < 2 * npts; r += nth) {
                const int rc = r >= npts, p = r - rc * npts;
                const int i = p % g.Nx + 1, j = p / g.Nx + 1;
                int I1, I2, J1, J2;
                T wx1, wx2, wy1, wy2;
                parents<T>(i, cg.Nx, g.tx, rc ? 0 : 1, C.hx, I1, I2, wx1, wx2);
                parents<T>(j, cg.Ny, g.ty, rc ? 1 : 0, C.hy, J1, J2, wy1, wy2);
                auto e = [&](int I, int J) { return rc ? iv(cg, cact, cv, I, J) : iu(cg, cact, cu, I, J); };
                const T pr = wx1 * (wy1 * e(I1, J1) + wy2 * e(I1, J2)) + wx2 * (wy1 * e(I2, J1) + wy2 * e(I2, J2));
                const long o = off(g, i, j);
                (rc ? xv : xu)[o] += (fact[o] & (rc ? 2 : 1)) ? pr : T(0);
            }
        }
        __syncthreads();
        NK_TT(nt++);
    }
    const GTailLevel<T> &L = d.lv[0];
    const T *xu = X(0, C_(0), 0), *xv = X(0, C_(0), 1);
    for (int p = tid; p < L.lg.Nx * L.lg.Ny; p += nth) {
        const int i = p % L.lg.Nx, j = p / L.lg.Nx;
        d.gxu[(long)j * L.gpitch + i] = xu[j * L.lg.pitch + i];
        d.gxv[(long)j * L.gpitch + i] = xv[j * L.lg.pitch + i];
    }
#ifdef NK_TAIL_TIMING
    __syncthreads();
    NK_TT(nt++);
    if (threadIdx.x == 0 && atomicAdd(&nk_tail_calls, 1u) == 40u) {
        printf("tail n=%d cinv=%d sweeps=%d coarsest=%d\n", d.n, d.cinv != nullptr, d.sweeps, d.coarsest);
        for (int k = 0; k < d.n; k++) printf("  level %d: %d x %d stencil %s\n", k, d.lv[k].lg.Nx, d.lv[k].lg.Ny, d.lv[k].o_st >= 0 ? "LDS" : "global");
        for (int k = 1; k < nt && k < 64; k++) printf("  stamp %d: +%.2f us\n", k, (double)(tt[k] - tt[k - 1]) * 0.01);
    }
#endif
}

}  // namespace nk
