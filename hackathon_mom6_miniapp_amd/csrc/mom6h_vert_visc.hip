// mom6h_vert_visc.hip -- vertical viscosity on gfx950.
//
// vert_visc_coef (src/serial/mom6_vert_visc_serial.F90:138-287) and
// vert_visc_apply (:394-506) fused into one kernel, one thread per velocity
// column (u columns I = is-1..ie, j = js..je and v columns i = is..ie,
// J = js-1..je), k-strided loads coalesced along i.  Three sweeps per column:
//   A (bottom-up)  harmonic-mean thickness, upwind blend and z_i;
//                  hvel parks in h_u(k), z_i(k) in a(k)
//   B (top-down)   coupling a(K) (consuming z_i(K)) fused with the forward
//                  elimination of the Schopf & Loughe tridiagonal system
//   C (bottom-up)  back substitution and the bottom stress
// Rayleigh drag is a 2D bottom-layer coefficient: the reference's Ray(k) is
// zero above the bottom layer (rk2_serial_driver.F90:524-540), so the same
// expressions are evaluated with Ray = 0 there and the bottom-stress sum keeps
// only its bottom term.
#include <type_traits>

#include "mom6h_internal.h"

namespace mom6h {

__device__ inline double botfn_of(double z2)
{
    return 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
}

struct VVCol {
    const double *hA, *hB;   // thickness columns at k=1 (h(i) and its neighbour)
    double *vel, *a, *hu, *c1;
    long long s;             // k stride
};

__device__ void vv_column(const VVCol &c, int nk, double mask, double dt, double sfc_stress,
                          double Ray_bot, double Kv, double Kv_ml, double Kv_extra_bbl, double Hmix,
                          double Hbbl, double *tbot)
{
    const double h_neglect = 1.0e-10;
    const long long s = c.s;
    if (!(mask > 0.0)) {  // masked column: coefficients only (:213-221)
        for (int k = 1; k <= nk; ++k) c.hu[(k - 1) * s] = h_neglect;
        for (int K = 1; K <= nk + 1; ++K) c.a[(K - 1) * s] = 0.0;
        return;
    }
    const double I_Hbbl = 1.0 / (Hbbl + h_neglect);
    // A: bottom-up (:160-179)
    double zi = 0.0;  // z_i(k+1)
    for (int k = nk; k >= 1; --k) {
        const long long o = (long long)(k - 1) * s;
        const double ha = c.hA[o], hb = c.hB[o];
        const double h_harm = 2.0 * ha * hb / (ha + hb + h_neglect);
        const double h_arith = 0.5 * (hb + ha);
        const double h_delta = hb - ha;
        double hv = h_harm;
        if (c.vel[o] * h_delta < 0.0) {
            const double botfn = botfn_of(zi);
            hv = (1.0 - botfn) * h_harm + botfn * h_arith;
        }
        c.hu[o] = hv;             // raw hvel(k)
        zi = zi + h_harm * I_Hbbl;
        c.a[o] = zi;              // z_i(k)
    }
    // a(K2) for K2 = 2..nk (:184-202) and a(nk+1) (:205-206)
    auto coefA = [&](int K2, double hv_m1, double hv_0, double &z_top) -> double {
        z_top = z_top + hv_m1;
        double Kv_tot = Kv;
        if (z_top < Hmix) {
            const double topfn = 1.0 - z_top / Hmix;
            Kv_tot = Kv_tot + (Kv_ml - Kv) * topfn;
        }
        const double botfn = botfn_of(c.a[(long long)(K2 - 1) * s]);  // z_i(K2)
        Kv_tot = Kv_tot + Kv_extra_bbl * botfn;
        const double h_shear = 0.5 * (hv_0 + hv_m1 + h_neglect);
        return Kv_tot / h_shear;
    };
    // B: top-down coefficients + forward elimination (:430-444)
    double z_top = 0.0;
    double hv_k = c.hu[0];
    double a_k = 0.0;  // a(1)
    double hv_n = (nk >= 2) ? c.hu[s] : 0.0;
    double a_n = (nk >= 2) ? coefA(2, hv_k, hv_n, z_top) : (Kv + Kv_extra_bbl) / (0.5 * hv_k + h_neglect);
    c.a[0] = 0.0;
    double hu_k = hv_k + h_neglect;
    c.hu[0] = hu_k;
    double Ray = (nk == 1) ? Ray_bot : 0.0;
    double b_denom_1 = hu_k + dt * (Ray + a_k);
    double b1 = 1.0 / (b_denom_1 + dt * a_n);
    double d1 = b_denom_1 * b1;
    double vprev = b1 * (hu_k * c.vel[0] + sfc_stress);
    c.vel[0] = vprev;
    for (int k = 2; k <= nk; ++k) {
        const long long o = (long long)(k - 1) * s;
        a_k = a_n;
        c.a[o] = a_k;
        hv_k = hv_n;
        if (k < nk) {
            hv_n = c.hu[o + s];
            a_n = coefA(k + 1, hv_k, hv_n, z_top);
        } else {
            a_n = (Kv + Kv_extra_bbl) / (0.5 * hv_k + h_neglect);
        }
        hu_k = hv_k + h_neglect;
        c.hu[o] = hu_k;
        Ray = (k == nk) ? Ray_bot : 0.0;
        c.c1[o] = dt * a_k * b1;
        b_denom_1 = hu_k + dt * (Ray + a_k * d1);
        b1 = 1.0 / (b_denom_1 + dt * a_n);
        d1 = b_denom_1 * b1;
        vprev = (hu_k * c.vel[o] + dt * a_k * vprev) * b1;
        c.vel[o] = vprev;
    }
    c.a[(long long)nk * s] = a_n;
    // C: back substitution (:447-449) and bottom stress (:452-458)
    double vnext = vprev;  // vel(nk)
    for (int k = nk - 1; k >= 1; --k) {
        const long long o = (long long)(k - 1) * s;
        vnext = c.vel[o] + c.c1[o + s] * vnext;
        c.vel[o] = vnext;
    }
    const double RHO_0 = 1035.0;
    const double vbot = c.vel[(long long)(nk - 1) * s];
    double t = RHO_0 * vbot * a_n;
    t = t + RHO_0 * Ray_bot * vbot;
    *tbot = t;
}

__global__ void k_vert_visc(Dims d, double *__restrict__ u, double *__restrict__ v,
                            const double *__restrict__ h, double dt, const double *__restrict__ taux,
                            const double *__restrict__ tauy, const double *__restrict__ ray_u,
                            const double *__restrict__ ray_v, const double *__restrict__ mask2dCu,
                            const double *__restrict__ mask2dCv, double *__restrict__ a_u,
                            double *__restrict__ a_v, double *__restrict__ h_u,
                            double *__restrict__ h_v, double *__restrict__ c1u,
                            double *__restrict__ c1v, double *__restrict__ taux_bot,
                            double *__restrict__ tauy_bot, double Kv, double Kv_ml,
                            double Kv_extra_bbl, double Hmix, double Hbbl)
{
    const int rows_u = d.jec - d.jsc + 1;
    const int row = blockIdx.y;
    const double dt_Rho0 = dt / 1035.0;
    VVCol c;
    c.s = d.s2;
    if (row < rows_u) {  // u column (I, j)
        const int I = d.isc - 1 + blockIdx.x * blockDim.x + threadIdx.x;
        const int j = d.jsc + row;
        if (I > d.iec) return;
        const long long c2 = IX2(d, I, j), o = IX3(d, I, j, 1);
        c.hA = h + o; c.hB = h + IX3(d, I + 1, j, 1);
        c.vel = u + o; c.a = a_u + o; c.hu = h_u + o; c.c1 = c1u + o;
        const double m = mask2dCu[c2];
        const double sfc = taux ? dt_Rho0 * taux[c2] * m : 0.0;
        double tb = 0.0;
        vv_column(c, d.nk, m, dt, sfc, ray_u ? ray_u[c2] : 0.0, Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl, &tb);
        if (m > 0.0) taux_bot[c2] = tb;
    } else {             // v column (i, J)
        const int i = d.isc + blockIdx.x * blockDim.x + threadIdx.x;
        const int J = d.jsc - 1 + (row - rows_u);
        if (i > d.iec) return;
        const long long c2 = IX2(d, i, J), o = IX3(d, i, J, 1);
        c.hA = h + o; c.hB = h + IX3(d, i, J + 1, 1);
        c.vel = v + o; c.a = a_v + o; c.hu = h_v + o; c.c1 = c1v + o;
        const double m = mask2dCv[c2];
        const double sfc = tauy ? dt_Rho0 * tauy[c2] * m : 0.0;
        double tb = 0.0;
        vv_column(c, d.nk, m, dt, sfc, ray_v ? ray_v[c2] : 0.0, Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl, &tb);
        if (m > 0.0) tauy_bot[c2] = tb;
    }
}


// Chunked variant.  One thread per column; the only per-column scratch is one
// LDS slot per level, holding z_i(k) after sweep A and c1(k) after sweep B
// (slot nk+1 holds z_i(nk+1) = 0).  Sweep B re-derives hvel from h and the input
// velocity (reloaded) and leaves the forward-eliminated velocity in the
// velocity array, which sweep C reads back.
//
// Levels go in chunks (VV_A levels in sweeps A and C, VV_C in sweep B) with
// straight-line (branch-free) bodies: the
// loads of the next chunk are issued before the current one is computed, and
// inside a chunk every division that is off the tridiagonal recurrence (h_harm,
// botfn, the mixed-layer taper, the coupling coefficient) is independent across
// levels, so the compiler interleaves them.  Only the Thomas recurrence itself
// (one division per level) is serial.  Levels that do not fill a chunk take a
// one-level loop with the same arithmetic (a partial chunk guarded level by
// level measured slower: the guards split the body and stop the interleaving).
// The LDS column holds the kernel at one wave per SIMD, so the registers of
// the prefetch are free: the FP64 bodies use AGPRs beyond 256 VGPRs rather
// than scratch.  At nk = 75 these chunk sizes leave a tail of 1-3 levels per
// sweep; 8 / 16 measured 7-10% slower (an 11-level unprefetched tail in
// sweeps A and C), 16 / 32 and larger slower again.

constexpr int VV_C = 12;   // levels per chunk of sweep B (heavy bodies)
constexpr int VV_A = 24;   // levels per chunk of sweeps A and C (light bodies, latency-bound loads)
constexpr int VV_T = 4;    // levels after the last whole chunk of sweeps A and C loaded with that chunk

// LDS slot type of the column (z_i, then c1): FP64 for FP64 state; FP32 in the
// mixed-precision form, where z_i is already FP32 and c1 is rounded once to
// FP32 (the same rounding the forward-eliminated velocity gets when stored as
// FP32), which halves the column and gives two waves per SIMD.
template <typename T>
using VVSlot = std::conditional_t<sizeof(T) == 4, float, double>;

// Block (bx, row) of one grid: 64 columns of one row; Zs = (nk+1)*64 slots of LDS.
// Rows are interleaved v, u, v, u, ... (v row J = js-1+r/2, u row j = js+(r-1)/2)
// so that consecutive rows share their h rows.
// RAY3: Rayleigh drag Ray(k) at every level from the 3D arrays ray_u3/ray_v3
// (vert_visc_apply :427-458 with a general visc%Ray_u); otherwise the 2D
// bottom-layer coefficient ray_u/ray_v and zero above it.
template <typename T, bool KEEP, bool RAY3 = false>
__device__ __forceinline__ void vv_body(const VVArgs<T> &p, int bx, int row, VVSlot<T> *Zs)
{
    const Dims &d = p.d;
    const int lane = threadIdx.x;
    const bool is_u = (row & 1) != 0;
    const long long s = d.s2;
    const int nk = d.nk;
    const T *hA, *hB;
    T *vel;
    double *a_out, *hu_out, *tb_out;
    const double *ray3 = nullptr;   // this column's Ray(1..nk) (RAY3)
    double m, sfc, Ray_bot;
    {
        const double dt_Rho0 = p.dt / 1035.0;
        if (is_u) {   // u column (I, j), I = is-1..ie
            const int I = d.isc - 1 + bx * 64 + lane, j = d.jsc + (row - 1) / 2;
            if (I > d.iec) return;
            const long long c2 = IX2(d, I, j), o = IX3(d, I, j, 1);
            hA = p.h + o; hB = p.h + IX3(d, I + 1, j, 1); vel = p.u + o;
            a_out = p.a_u + o; hu_out = p.h_u + o;
            tb_out = p.taux_bot + c2;
            m = p.mask2dCu[c2];
            sfc = p.taux ? dt_Rho0 * p.taux[c2] * m : 0.0;
            Ray_bot = p.ray_u ? p.ray_u[c2] : 0.0;
            if (RAY3) ray3 = p.ray_u3 + o;
        } else {              // v column (i, J), J = js-1..je
            const int i = d.isc + bx * 64 + lane, J = d.jsc - 1 + row / 2;
            if (i > d.iec) return;
            const long long c2 = IX2(d, i, J), o = IX3(d, i, J, 1);
            hA = p.h + o; hB = p.h + IX3(d, i, J + 1, 1); vel = p.v + o;
            a_out = p.a_v + o; hu_out = p.h_v + o;
            tb_out = p.tauy_bot + c2;
            m = p.mask2dCv[c2];
            sfc = p.tauy ? dt_Rho0 * p.tauy[c2] * m : 0.0;
            Ray_bot = p.ray_v ? p.ray_v[c2] : 0.0;
            if (RAY3) ray3 = p.ray_v3 + o;
        }
    }
    const double h_neglect = 1.0e-10;
    // A masked column (:213-221) keeps its velocity and takes the coefficients
    // hu = h_neglect and a = 0.  It runs the sweeps like the others (on its
    // own in-bounds data) and its stores write those values (selects, no
    // branch), so that no load waits for the mask: returning early would hold
    // every column's first loads behind the mask's memory round trip.
    const bool act = m > 0.0;
    using ZT = VVSlot<T>;
    auto Z = [&](int k) -> ZT & { return Zs[(k - 1) * 64 + lane]; };   // level k = 1..nk+1
    // C: the arithmetic type of the coefficients (z_i, bottom taper, hvel, the
    // coupling a); the tridiagonal solve is FP64 whatever C is
    using C = T;
    // division of the coefficients: exact in FP64; in the FP32 form a product
    // with the hardware reciprocal (about 1 ulp, the rounding scale of the
    // FP32 coefficients themselves; the tridiagonal solve stays FP64)
    auto cdiv = [](C num, C den) __attribute__((always_inline)) -> C {
        if constexpr (sizeof(C) == 4) return num * __builtin_amdgcn_rcpf(den);
        else return num / den;
    };
    const C hn = (C)h_neglect;
    const C Kv = (C)p.Kv, Kv_ml = (C)p.Kv_ml, Kv_extra_bbl = (C)p.Kv_extra_bbl, Hmix = (C)p.Hmix;
    const double dt = p.dt;
    const C I_Hbbl = C(1.0) / ((C)p.Hbbl + hn);
    auto botfn = [&](C z2) -> C { return cdiv(C(1.0), C(1.0) + C(0.09) * z2 * z2 * z2 * z2 * z2 * z2); };
    auto ZC = [&](int k) -> C { return (C)Z(k); };
    Z(nk + 1) = 0.0;

    // ---- sweep A, bottom-up (:160-179): z_i(k) ----
    C zi = C(0.0);
    auto A_level = [&](int k, C ha, C hb) {
        const C h_harm = cdiv(C(2.0) * ha * hb, ha + hb + hn);
        zi = zi + h_harm * I_Hbbl;
        Z(k) = (ZT)zi;
    };
    int k0 = nk;
    C ta[VV_T], tb[VV_T];   // the first VV_T levels after the last whole chunk (clamped to level 1)
    bool tail_ready = false;
    if (k0 >= VV_A) {
        C na[VV_A], nb[VV_A];
#pragma unroll
        for (int c = 0; c < VV_A; ++c) { na[c] = hA[(k0 - 1 - c) * s]; nb[c] = hB[(k0 - 1 - c) * s]; }
        for (; k0 >= VV_A; k0 -= VV_A) {   // levels k0 .. k0-VV_A+1, all >= 1
            C ha[VV_A], hb[VV_A];
#pragma unroll
            for (int c = 0; c < VV_A; ++c) { ha[c] = na[c]; hb[c] = nb[c]; }
            if (k0 - VV_A >= VV_A) {
#pragma unroll
                for (int c = 0; c < VV_A; ++c) {
                    na[c] = hA[(k0 - VV_A - 1 - c) * s];
                    nb[c] = hB[(k0 - VV_A - 1 - c) * s];
                }
            } else if (k0 - VV_A >= 1) {
#pragma unroll
                for (int c = 0; c < VV_T; ++c) {
                    const int k = k0 - VV_A - c >= 1 ? k0 - VV_A - c : 1;
                    ta[c] = hA[(k - 1) * s];
                    tb[c] = hB[(k - 1) * s];
                }
                tail_ready = true;
            }
#pragma unroll
            for (int c = 0; c < VV_A; ++c) A_level(k0 - c, ha[c], hb[c]);
        }
    }
    // sweep B's first loads (levels 1 and 2, the first chunk's level k+1
    // terms), issued before the last levels of sweep A are folded; sweep A
    // writes no velocity and sweep B stores none of these levels before it reads them
    const C p_v1 = vel[0], p_a1 = hA[0], p_b1 = hB[0];
    C p_v2 = C(0.0), p_a2 = C(0.0), p_b2 = C(0.0);
    if (nk >= 2) { p_v2 = vel[s]; p_a2 = hA[s]; p_b2 = hB[s]; }
    C na[VV_C], nb[VV_C], nv[VV_C];
    const bool b_chunks = 2 + VV_C - 1 <= nk - 1;
    if (b_chunks) {
#pragma unroll
        for (int c = 0; c < VV_C; ++c) {   // h and u at k+1 for the chunk's levels k = 2 .. VV_C+1
            na[c] = hA[(2 + c) * s]; nb[c] = hB[(2 + c) * s]; nv[c] = vel[(2 + c) * s];
        }
    }
    if (tail_ready) {
#pragma unroll
        for (int c = 0; c < VV_T; ++c)
            if (k0 - c >= 1) A_level(k0 - c, ta[c], tb[c]);
        k0 -= VV_T;
    }
    for (; k0 >= 1; --k0) A_level(k0, hA[(k0 - 1) * s], hB[(k0 - 1) * s]);

    // hvel (:165-176) given botfn(z_i(k+1)); the upwind blend is selected, not branched
    auto hvel = [&](C ha, C hb, C vk, C bf) -> C {
        const C h_harm = cdiv(C(2.0) * ha * hb, ha + hb + hn);
        const C h_arith = C(0.5) * (hb + ha);
        const C h_delta = hb - ha;
        const C blend = (C(1.0) - bf) * h_harm + bf * h_arith;
        return (vk * h_delta < C(0.0)) ? blend : h_harm;
    };
    // Kv with the mixed-layer taper at z_top (:188-192).  The tapered value
    // passes through an empty asm so that the compiler keeps the select: with
    // a division among its operands it turns the select into a branch, one per
    // level, which splits a chunk into a basic block per level and stops the
    // interleaving of the levels' independent divisions.
    auto kv_top = [&](C z_top) __attribute__((always_inline)) -> C {
        C t = Kv + (Kv_ml - Kv) * (C(1.0) - z_top / Hmix);
        asm("" : "+v"(t));
        return (z_top < Hmix) ? t : Kv;
    };
    // coupling coefficient a(K2) (:184-202) from kv_top(z_top) (z_top
    // already advanced), botfn(z_i(K2)) and hvel at K2-1 and K2
    auto coefA = [&](C kvt, C bfK, C hv_m1, C hv_0) -> C {
        const C Kv_tot = kvt + Kv_extra_bbl * bfK;
        return cdiv(Kv_tot, C(0.5) * (hv_0 + hv_m1 + hn));
    };

    // ---- sweep B, top-down (:430-444) ----
    C z_top, hv_n;
    double a_n, vk_cur, b1, d1, vprev;
    {   // level 1
        const C v1 = p_v1;
        const C hv1 = hvel(p_a1, p_b1, v1, botfn(ZC(2)));
        hv_n = C(0.0);
        z_top = C(0.0);
        if (nk >= 2) {
            const C v2 = p_v2;
            hv_n = hvel(p_a2, p_b2, v2, botfn(ZC(3)));
            vk_cur = (double)v2;
            z_top = z_top + hv1;
            a_n = (double)coefA(kv_top(z_top), botfn(ZC(2)), hv1, hv_n);
        } else {
            vk_cur = 0.0;
            a_n = (double)((Kv + Kv_extra_bbl) / (C(0.5) * hv1 + hn));
        }
        if (KEEP) a_out[0] = 0.0;
        const double hu_k = (double)hv1 + h_neglect;
        if (KEEP) hu_out[0] = act ? hu_k : h_neglect;
        const double Ray = RAY3 ? ray3[0] : ((nk == 1) ? Ray_bot : 0.0);
        const double b_denom_1 = hu_k + dt * (Ray + 0.0);
        b1 = 1.0 / (b_denom_1 + dt * a_n);
        d1 = b_denom_1 * b1;
        vprev = b1 * (hu_k * (double)v1 + sfc);
        vel[0] = act ? (T)vprev : v1;
    }
    // one interior level k (2..nk-1): the recurrence, given a(k+1) and hvel(k+1)
    auto B_serial = [&](int k, double a_next, C hv_next, double vk_next) {
        const double a_k = a_n;
        if (KEEP) a_out[(k - 1) * s] = act ? a_k : 0.0;
        const C hv_k = hv_n;
        const double hu_k = (double)hv_k + h_neglect;
        if (KEEP) hu_out[(k - 1) * s] = act ? hu_k : h_neglect;
        Z(k) = (ZT)(dt * a_k * b1);                        // c1(k); z_i(k) is no longer read
        const double b_denom_1 = hu_k + dt * ((RAY3 ? ray3[(k - 1) * s] : 0.0) + a_k * d1);
        b1 = 1.0 / (b_denom_1 + dt * a_next);
        d1 = b_denom_1 * b1;
        vprev = (hu_k * vk_cur + dt * a_k * vprev) * b1;
        vel[(k - 1) * s] = act ? (T)vprev : (T)vk_cur;
        a_n = a_next;
        hv_n = hv_next;
        vk_cur = vk_next;
    };
    k0 = 2;
    C tb_a = C(0.0), tb_b = C(0.0), tb_v = C(0.0);   // the first level after the last whole chunk
    bool tb_ready = false;
    if (b_chunks) {
        for (; k0 + VV_C - 1 <= nk - 1; k0 += VV_C) {   // levels k0 .. k0+VV_C-1
            C ha[VV_C], hb[VV_C], vv[VV_C];
#pragma unroll
            for (int c = 0; c < VV_C; ++c) { ha[c] = na[c]; hb[c] = nb[c]; vv[c] = nv[c]; }
            if (k0 + 2 * VV_C - 1 <= nk - 1) {
#pragma unroll
                for (int c = 0; c < VV_C; ++c) {
                    na[c] = hA[(k0 + VV_C + c) * s];
                    nb[c] = hB[(k0 + VV_C + c) * s];
                    nv[c] = vel[(k0 + VV_C + c) * s];
                }
            } else if (k0 + VV_C <= nk - 1) {   // its velocity is stored after this load
                tb_a = hA[(k0 + VV_C) * s];
                tb_b = hB[(k0 + VV_C) * s];
                tb_v = vel[(k0 + VV_C) * s];
                tb_ready = true;
            }
            // off-recurrence terms: bf[c] = botfn(z_i(k0+1+c)), hvel(k+1), z_top, a(k+1)
            C bf[VV_C + 1], hvx[VV_C], ax[VV_C], hvk[VV_C], zt[VV_C], kvt[VV_C];
#pragma unroll
            for (int c = 0; c <= VV_C; ++c) bf[c] = botfn(ZC(k0 + 1 + c));
#pragma unroll
            for (int c = 0; c < VV_C; ++c) hvx[c] = hvel(ha[c], hb[c], vv[c], bf[c + 1]);
            bool taper = false;
#pragma unroll
            for (int c = 0; c < VV_C; ++c) {
                hvk[c] = (c == 0) ? hv_n : hvx[c - 1];
                z_top = z_top + hvk[c];
                zt[c] = z_top;
                taper = taper || (z_top < Hmix);
            }
            // the taper only in a chunk where some column has a level above
            // Hmix (kv_top is Kv at every other level).  Its inputs pass
            // through a volatile asm inside the block, so that the divisions
            // stay behind the one scalar branch per chunk instead of being
            // hoisted above it and evaluated in every chunk.
            if (__any(taper)) {
#pragma unroll
                for (int c = 0; c < VV_C; ++c) {
                    C z = zt[c];
                    asm volatile("" : "+v"(z));
                    kvt[c] = kv_top(z);
                }
            } else {
#pragma unroll
                for (int c = 0; c < VV_C; ++c) kvt[c] = Kv;
            }
#pragma unroll
            for (int c = 0; c < VV_C; ++c) ax[c] = coefA(kvt[c], bf[c], hvk[c], hvx[c]);
#pragma unroll
            for (int c = 0; c < VV_C; ++c) B_serial(k0 + c, (double)ax[c], hvx[c], (double)vv[c]);
        }
    }
    for (; k0 <= nk - 1; ++k0) {
        C vnx, ha_, hb_;
        if (tb_ready) {
            vnx = tb_v; ha_ = tb_a; hb_ = tb_b;
            tb_ready = false;
        } else {
            vnx = vel[k0 * s]; ha_ = hA[k0 * s]; hb_ = hB[k0 * s];
        }
        const C hvx = hvel(ha_, hb_, vnx, botfn(ZC(k0 + 2)));
        z_top = z_top + hv_n;
        const C ax = coefA(kv_top(z_top), botfn(ZC(k0 + 1)), hv_n, hvx);
        B_serial(k0, (double)ax, hvx, (double)vnx);
    }
    if (nk >= 2) {   // bottom level: a(nk+1) from the bottom formula (:205-206), Rayleigh drag
        const int k = nk;
        const double a_k = a_n;
        if (KEEP) a_out[(k - 1) * s] = act ? a_k : 0.0;
        const C hv_k = hv_n;
        a_n = (double)((Kv + Kv_extra_bbl) / (C(0.5) * hv_k + hn));
        const double hu_k = (double)hv_k + h_neglect;
        if (KEEP) hu_out[(k - 1) * s] = act ? hu_k : h_neglect;
        Z(k) = (ZT)(dt * a_k * b1);
        const double b_denom_1 = hu_k + dt * ((RAY3 ? ray3[(k - 1) * s] : Ray_bot) + a_k * d1);
        b1 = 1.0 / (b_denom_1 + dt * a_n);
        d1 = b_denom_1 * b1;
        vprev = (hu_k * vk_cur + dt * a_k * vprev) * b1;
        vel[(k - 1) * s] = act ? (T)vprev : (T)vk_cur;
    }
    if (KEEP) a_out[nk * s] = act ? a_n : 0.0;

    // ---- sweep C, bottom-up back substitution (:447-449); a masked column
    // stores back the velocity sweep B left in place (its input) ----
    double vnext = vprev;
    k0 = nk - 1;
    double tf[VV_T];   // the first VV_T levels after the last whole chunk (clamped to level 1)
    bool tc_ready = false;
    if (k0 >= VV_A) {
        double nf[VV_A];
#pragma unroll
        for (int c = 0; c < VV_A; ++c) nf[c] = vel[(k0 - 1 - c) * s];
        for (; k0 >= VV_A; k0 -= VV_A) {   // levels k0 .. k0-VV_A+1
            double fv[VV_A];
#pragma unroll
            for (int c = 0; c < VV_A; ++c) fv[c] = nf[c];
            if (k0 - VV_A >= VV_A) {
#pragma unroll
                for (int c = 0; c < VV_A; ++c) nf[c] = vel[(k0 - VV_A - 1 - c) * s];
            } else if (k0 - VV_A >= 1) {   // below this chunk's stores
#pragma unroll
                for (int c = 0; c < VV_T; ++c) {
                    const int k = k0 - VV_A - c >= 1 ? k0 - VV_A - c : 1;
                    tf[c] = vel[(k - 1) * s];
                }
                tc_ready = true;
            }
#pragma unroll
            for (int c = 0; c < VV_A; ++c) {
                vnext = fv[c] + Z(k0 - c + 1) * vnext;
                vel[(k0 - c - 1) * s] = act ? (T)vnext : (T)fv[c];
            }
        }
    }
    if (tc_ready) {
#pragma unroll
        for (int c = 0; c < VV_T; ++c) {
            if (k0 - c >= 1) {
                vnext = tf[c] + Z(k0 - c + 1) * vnext;
                vel[(k0 - c - 1) * s] = act ? (T)vnext : (T)tf[c];
            }
        }
        k0 -= VV_T;
    }
    for (; k0 >= 1; --k0) {
        const double f = vel[(k0 - 1) * s];
        vnext = f + Z(k0 + 1) * vnext;
        vel[(k0 - 1) * s] = act ? (T)vnext : (T)f;
    }
    // bottom stress (:452-458)
    const double RHO_0 = 1035.0;
    double t = RHO_0 * vprev * a_n;
    if (RAY3) {   // + sum_k RHO_0*Ray(k)*u(k) of the solved column, in k order
        for (int k = 1; k <= nk; ++k) t = t + RHO_0 * ray3[(k - 1) * s] * (double)vel[(k - 1) * s];
    } else {
        t = t + RHO_0 * Ray_bot * vprev;
    }
    if (act) *tb_out = t;
}

// one grid, tiles renumbered per XCD
template <typename T, bool KEEP, bool RAY3 = false>
// FP64: one wave per SIMD (the LDS column), registers up to 512 with AGPRs;
// FP32 state: the half-size column allows two, and the register budget keeps them
__global__ void __launch_bounds__(64, 1) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
k_vert_visc_chunked(VVArgs<T> p)
{
    extern __shared__ double Zs_raw[];
    VVSlot<T> *Zs = reinterpret_cast<VVSlot<T> *>(Zs_raw);
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    vv_body<T, KEEP, RAY3>(p, tile % gridDim.x, tile / gridDim.x, Zs);
}

// every member of an ensemble in one launch (LDS sized for the deepest member)
template <typename T>
__global__ void __launch_bounds__(64, 1) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
k_vert_visc_batch(const VVArgs<T> *__restrict__ args, BatchMap mp)
{
    extern __shared__ double Zs_raw[];
    VVSlot<T> *Zs = reinterpret_cast<VVSlot<T> *>(Zs_raw);
    int m, local;
    batch_locate(mp, xcd_tile(blockIdx.x, gridDim.x), m, local);
    const VVArgs<T> p = args[m];
    vv_body<T, false>(p, local % mp.nbx[m], local / mp.nbx[m], Zs);
}

// The mixed-precision form: FP32 u, v, h in memory, FP64 solve; the
// coefficients stay on chip (the LDS kernel only).
int vert_visc_f32(mom6h_vert_visc_cs *cs, float *u, float *v, const float *h, double dt, const double *taux,
                  const double *tauy, const double *ray_u_bot, const double *ray_v_bot, hipStream_t stream)
{
    const Dims &d = cs->d;
    const int ni = d.iec - d.isc + 2;
    const int rows = (d.jec - d.jsc + 1) + (d.jec - d.jsc + 2);
    const size_t lds = (size_t)(d.nk + 1) * 64 * sizeof(VVSlot<float>);
    MOM6H_REQUIRE((long long)lds <= cs->max_lds, "vert_visc (FP32 state): nk = %d exceeds the LDS column (%lld B)",
                  d.nk, cs->max_lds);
    VVArgs<float> p{d, u, v, h, dt, taux, tauy, ray_u_bot, ray_v_bot, cs->mask2dCu, cs->mask2dCv,
                    cs->a_u, cs->a_v, cs->h_u, cs->h_v, cs->taux_bot, cs->tauy_bot, cs->Kv, cs->Kv_ml,
                    cs->Kv_extra_bbl, cs->Hmix, cs->Hbbl};
    k_vert_visc_chunked<float, false><<<dim3((ni + 63) / 64, rows), dim3(64), lds, stream>>>(p);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

template <typename T>
int vert_visc_batch(const VVArgs<T> *d_args, const VVArgs<T> *h_args, int n, long long max_lds, hipStream_t s)
{
    MOM6H_REQUIRE(n >= 1 && n <= kMaxMembers, "vert_visc batch: 1..%d members", kMaxMembers);
    BatchMap mp{};
    mp.n = n;
    int nk_max = 1;
    for (int m = 0; m < n; ++m) {
        const Dims &d = h_args[m].d;
        mp.nbx[m] = (d.iec - d.isc + 2 + 63) / 64;
        mp.nby[m] = (d.jec - d.jsc + 1) + (d.jec - d.jsc + 2);
        mp.first[m + 1] = mp.first[m] + mp.nbx[m] * mp.nby[m];
        nk_max = d.nk > nk_max ? d.nk : nk_max;
    }
    const size_t lds = (size_t)(nk_max + 1) * 64 * sizeof(VVSlot<T>);
    MOM6H_REQUIRE((long long)lds <= max_lds, "vert_visc batch: nk = %d exceeds the LDS column", nk_max);
    k_vert_visc_batch<T><<<(unsigned)mp.first[n], 64, lds, s>>>(d_args, mp);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

template int vert_visc_batch<float>(const VVArgs<float> *, const VVArgs<float> *, int, long long, hipStream_t);
template int vert_visc_batch<double>(const VVArgs<double> *, const VVArgs<double> *, int, long long, hipStream_t);

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_vert_visc_init(mom6h_vert_visc_cs **out, int isd, int ied, int jsd, int jed, int isc,
                         int iec, int jsc, int jec, int nk, const double *mask2dCu,
                         const double *mask2dCv, double Kv, double Kv_ml, double Kv_extra_bbl,
                         double Hmix, double Hbbl)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_vert_visc_init: NULL cs");
    MOM6H_REQUIRE(isd < isc && iec < ied && jsd < jsc && jec < jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_vert_visc_init: needs one halo cell on every side");
    MOM6H_REQUIRE(mask2dCu && mask2dCv, "mom6h_vert_visc_init: NULL mask");
    auto *cs = new mom6h_vert_visc_cs();
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    cs->Kv = Kv; cs->Kv_ml = Kv_ml; cs->Kv_extra_bbl = Kv_extra_bbl; cs->Hmix = Hmix; cs->Hbbl = Hbbl;
    cs->keep_coefs = 1;
    {
        int dev = 0, lds = 0;
        MOM6H_CHECK(hipGetDevice(&dev));
        MOM6H_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cs->max_lds = lds;
    }
    const size_t n2 = (size_t)cs->d.s2, n3 = n2 * nk;
    int rc = 0;
    rc = rc ? rc : dev_upload(&cs->mask2dCu, mask2dCu, n2);
    rc = rc ? rc : dev_upload(&cs->mask2dCv, mask2dCv, n2);
    rc = rc ? rc : dev_alloc(&cs->a_u, n3 + n2, true);
    rc = rc ? rc : dev_alloc(&cs->a_v, n3 + n2, true);
    rc = rc ? rc : dev_alloc(&cs->h_u, n3, true);
    rc = rc ? rc : dev_alloc(&cs->h_v, n3, true);
    rc = rc ? rc : dev_alloc(&cs->c1, 2 * n3, true);
    rc = rc ? rc : dev_alloc(&cs->taux_bot, n2, true);
    rc = rc ? rc : dev_alloc(&cs->tauy_bot, n2, true);
    if (rc) {
        mom6h_vert_visc_end(cs);
        return rc;
    }
    *out = cs;
    return MOM6H_OK;
}

int mom6h_vert_visc(mom6h_vert_visc_cs *cs, double *u, double *v, const double *h, double dt,
                    const double *taux, const double *tauy, const double *ray_u_bot,
                    const double *ray_v_bot, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_vert_visc: NULL cs");
    MOM6H_REQUIRE(u && v && h, "mom6h_vert_visc: NULL array");
    const Dims &d = cs->d;
    const int ni = d.iec - d.isc + 2;
    const int rows = (d.jec - d.jsc + 1) + (d.jec - d.jsc + 2);
    const size_t lds = (size_t)(d.nk + 1) * 64 * sizeof(double);
    if ((long long)lds <= cs->max_lds) {
        VVArgs<double> p{d, u, v, h, dt, taux, tauy, ray_u_bot, ray_v_bot, cs->mask2dCu, cs->mask2dCv,
                 cs->a_u, cs->a_v, cs->h_u, cs->h_v, cs->taux_bot, cs->tauy_bot, cs->Kv, cs->Kv_ml,
                 cs->Kv_extra_bbl, cs->Hmix, cs->Hbbl};
        const dim3 g((ni + 63) / 64, rows);
        if (cs->keep_coefs)
            k_vert_visc_chunked<double, true><<<g, dim3(64), lds, (hipStream_t)stream>>>(p);
        else
            k_vert_visc_chunked<double, false><<<g, dim3(64), lds, (hipStream_t)stream>>>(p);
        MOM6H_LAUNCH_CHECK();
        return MOM6H_OK;
    }
    dim3 b(64), g((ni + 63) / 64, rows);
    const size_t n3 = (size_t)d.s2 * d.nk;
    k_vert_visc<<<g, b, 0, (hipStream_t)stream>>>(d, u, v, h, dt, taux, tauy, ray_u_bot, ray_v_bot,
                                                  cs->mask2dCu, cs->mask2dCv, cs->a_u, cs->a_v, cs->h_u,
                                                  cs->h_v, cs->c1, cs->c1 + n3, cs->taux_bot,
                                                  cs->tauy_bot, cs->Kv, cs->Kv_ml, cs->Kv_extra_bbl,
                                                  cs->Hmix, cs->Hbbl);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_vert_visc_ray3d(mom6h_vert_visc_cs *cs, double *u, double *v, const double *h, double dt,
                          const double *taux, const double *tauy, const double *ray_u, const double *ray_v,
                          void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_vert_visc_ray3d: NULL cs");
    MOM6H_REQUIRE(u && v && h && ray_u && ray_v, "mom6h_vert_visc_ray3d: NULL array");
    const Dims &d = cs->d;
    const int ni = d.iec - d.isc + 2;
    const int rows = (d.jec - d.jsc + 1) + (d.jec - d.jsc + 2);
    const size_t lds = (size_t)(d.nk + 1) * 64 * sizeof(double);
    MOM6H_REQUIRE((long long)lds <= cs->max_lds, "mom6h_vert_visc_ray3d: nk = %d exceeds the LDS column (%lld B)",
                  d.nk, cs->max_lds);
    VVArgs<double> p{d, u, v, h, dt, taux, tauy, nullptr, nullptr, cs->mask2dCu, cs->mask2dCv,
                     cs->a_u, cs->a_v, cs->h_u, cs->h_v, cs->taux_bot, cs->tauy_bot, cs->Kv, cs->Kv_ml,
                     cs->Kv_extra_bbl, cs->Hmix, cs->Hbbl, ray_u, ray_v};
    const dim3 g((ni + 63) / 64, rows);
    if (cs->keep_coefs)
        k_vert_visc_chunked<double, true, true><<<g, dim3(64), lds, (hipStream_t)stream>>>(p);
    else
        k_vert_visc_chunked<double, false, true><<<g, dim3(64), lds, (hipStream_t)stream>>>(p);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

int mom6h_vert_visc_keep_coefficients(mom6h_vert_visc_cs *cs, int keep)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_vert_visc_keep_coefficients: NULL cs");
    cs->keep_coefs = keep;
    return MOM6H_OK;
}

int mom6h_vert_visc_arrays(mom6h_vert_visc_cs *cs, double **a_u, double **a_v, double **h_u,
                           double **h_v, double **taux_bot, double **tauy_bot)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_vert_visc_arrays: NULL cs");
    if (a_u) *a_u = cs->a_u;
    if (a_v) *a_v = cs->a_v;
    if (h_u) *h_u = cs->h_u;
    if (h_v) *h_v = cs->h_v;
    if (taux_bot) *taux_bot = cs->taux_bot;
    if (tauy_bot) *tauy_bot = cs->tauy_bot;
    return MOM6H_OK;
}

int mom6h_vert_visc_end(mom6h_vert_visc_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->mask2dCu, cs->mask2dCv, cs->a_u, cs->a_v, cs->h_u, cs->h_v, cs->c1,
                      cs->taux_bot, cs->tauy_bot})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
