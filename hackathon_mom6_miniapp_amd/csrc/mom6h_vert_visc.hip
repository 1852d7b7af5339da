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


// LDS variant: the per-column scratch lives in shared memory, one slot per
// (level, lane) so a wavefront's accesses are conflict-free:
//   SA[k]  z_i(k) during sweep A, then c1(k) during sweep B
//   SH[k]  hvel(k) during sweep A, then the forward-eliminated vel(k)
// HBM traffic per column is h(i), h(i+1), vel in and vel out; the coupling
// coefficients and effective thicknesses are written only when requested.
constexpr int VV_BLOCK = 64;

__device__ void vv_column_lds(double *SA, double *SH, int lane, const double *hA, const double *hB,
                              double *vel, double *a_out, double *hu_out, long long s, int nk,
                              double mask, double dt, double sfc_stress, double Ray_bot, double Kv,
                              double Kv_ml, double Kv_extra_bbl, double Hmix, double Hbbl, double *tbot)
{
    const double h_neglect = 1.0e-10;
    auto A = [&](int k) -> double & { return SA[(k - 1) * VV_BLOCK + lane]; };   // k = 1..nk+1
    auto H = [&](int k) -> double & { return SH[(k - 1) * VV_BLOCK + lane]; };
    if (!(mask > 0.0)) {  // masked column (:213-221)
        if (hu_out)
            for (int k = 1; k <= nk; ++k) hu_out[(long long)(k - 1) * s] = h_neglect;
        if (a_out)
            for (int K = 1; K <= nk + 1; ++K) a_out[(long long)(K - 1) * s] = 0.0;
        return;
    }
    const double I_Hbbl = 1.0 / (Hbbl + h_neglect);
    double zi = 0.0;
    for (int k = nk; k >= 1; --k) {   // sweep A (:160-179)
        const long long o = (long long)(k - 1) * s;
        const double ha = hA[o], hb = hB[o];
        const double h_harm = 2.0 * ha * hb / (ha + hb + h_neglect);
        const double h_arith = 0.5 * (hb + ha);
        const double h_delta = hb - ha;
        double hv = h_harm;
        if (vel[o] * h_delta < 0.0) {
            const double botfn = botfn_of(zi);
            hv = (1.0 - botfn) * h_harm + botfn * h_arith;
        }
        H(k) = hv;
        zi = zi + h_harm * I_Hbbl;
        A(k) = zi;
    }
    auto coefA = [&](int K2, double hv_m1, double hv_0, double &z_top) -> double {   // (:184-202)
        z_top = z_top + hv_m1;
        double Kv_tot = Kv;
        if (z_top < Hmix) {
            const double topfn = 1.0 - z_top / Hmix;
            Kv_tot = Kv_tot + (Kv_ml - Kv) * topfn;
        }
        const double botfn = botfn_of(A(K2));  // z_i(K2)
        Kv_tot = Kv_tot + Kv_extra_bbl * botfn;
        const double h_shear = 0.5 * (hv_0 + hv_m1 + h_neglect);
        return Kv_tot / h_shear;
    };
    // sweep B (:430-444) with the coefficients formed on the fly
    double z_top = 0.0;
    double hv_k = H(1);
    double a_k = 0.0;
    double hv_n = (nk >= 2) ? H(2) : 0.0;
    double a_n = (nk >= 2) ? coefA(2, hv_k, hv_n, z_top) : (Kv + Kv_extra_bbl) / (0.5 * hv_k + h_neglect);
    if (a_out) a_out[0] = 0.0;
    double hu_k = hv_k + h_neglect;
    if (hu_out) hu_out[0] = hu_k;
    double Ray = (nk == 1) ? Ray_bot : 0.0;
    double b_denom_1 = hu_k + dt * (Ray + a_k);
    double b1 = 1.0 / (b_denom_1 + dt * a_n);
    double d1 = b_denom_1 * b1;
    double vprev = b1 * (hu_k * vel[0] + sfc_stress);
    H(1) = vprev;
    for (int k = 2; k <= nk; ++k) {
        const long long o = (long long)(k - 1) * s;
        a_k = a_n;
        if (a_out) a_out[o] = a_k;
        hv_k = hv_n;
        if (k < nk) {
            hv_n = H(k + 1);
            a_n = coefA(k + 1, hv_k, hv_n, z_top);
        } else {
            a_n = (Kv + Kv_extra_bbl) / (0.5 * hv_k + h_neglect);
        }
        hu_k = hv_k + h_neglect;
        if (hu_out) hu_out[o] = hu_k;
        Ray = (k == nk) ? Ray_bot : 0.0;
        A(k) = dt * a_k * b1;            // c1(k); z_i(k) was consumed one level up
        b_denom_1 = hu_k + dt * (Ray + a_k * d1);
        b1 = 1.0 / (b_denom_1 + dt * a_n);
        d1 = b_denom_1 * b1;
        vprev = (hu_k * vel[o] + dt * a_k * vprev) * b1;
        H(k) = vprev;                    // hvel(k) was consumed one level up
    }
    if (a_out) a_out[(long long)nk * s] = a_n;
    // sweep C (:447-458)
    double vnext = vprev;
    vel[(long long)(nk - 1) * s] = vnext;
    for (int k = nk - 1; k >= 1; --k) {
        vnext = H(k) + A(k + 1) * vnext;
        vel[(long long)(k - 1) * s] = vnext;
    }
    const double RHO_0 = 1035.0;
    double t = RHO_0 * vprev * a_n;
    t = t + RHO_0 * Ray_bot * vprev;
    *tbot = t;
}

__global__ void __launch_bounds__(VV_BLOCK) k_vert_visc_lds(
    Dims d, double *__restrict__ u, double *__restrict__ v, const double *__restrict__ h, double dt,
    const double *__restrict__ taux, const double *__restrict__ tauy, const double *__restrict__ ray_u,
    const double *__restrict__ ray_v, const double *__restrict__ mask2dCu, const double *__restrict__ mask2dCv,
    double *__restrict__ a_u, double *__restrict__ a_v, double *__restrict__ h_u, double *__restrict__ h_v,
    double *__restrict__ taux_bot, double *__restrict__ tauy_bot, double Kv, double Kv_ml,
    double Kv_extra_bbl, double Hmix, double Hbbl)
{
    extern __shared__ double smem[];
    double *SA = smem, *SH = smem + (size_t)(d.nk + 1) * VV_BLOCK;
    const int lane = threadIdx.x;
    const int rows_u = d.jec - d.jsc + 1;
    const int row = blockIdx.y;
    const double dt_Rho0 = dt / 1035.0;
    if (row < rows_u) {
        const int I = d.isc - 1 + blockIdx.x * VV_BLOCK + lane;
        const int j = d.jsc + row;
        if (I > d.iec) return;
        const long long c2 = IX2(d, I, j), o = IX3(d, I, j, 1);
        const double m = mask2dCu[c2];
        const double sfc = taux ? dt_Rho0 * taux[c2] * m : 0.0;
        double tb = 0.0;
        vv_column_lds(SA, SH, lane, h + o, h + IX3(d, I + 1, j, 1), u + o, a_u ? a_u + o : nullptr,
                      h_u ? h_u + o : nullptr, d.s2, d.nk, m, dt, sfc, ray_u ? ray_u[c2] : 0.0, Kv, Kv_ml,
                      Kv_extra_bbl, Hmix, Hbbl, &tb);
        if (m > 0.0) taux_bot[c2] = tb;
    } else {
        const int i = d.isc + blockIdx.x * VV_BLOCK + lane;
        const int J = d.jsc - 1 + (row - rows_u);
        if (i > d.iec) return;
        const long long c2 = IX2(d, i, J), o = IX3(d, i, J, 1);
        const double m = mask2dCv[c2];
        const double sfc = tauy ? dt_Rho0 * tauy[c2] * m : 0.0;
        double tb = 0.0;
        vv_column_lds(SA, SH, lane, h + o, h + IX3(d, i, J + 1, 1), v + o, a_v ? a_v + o : nullptr,
                      h_v ? h_v + o : nullptr, d.s2, d.nk, m, dt, sfc, ray_v ? ray_v[c2] : 0.0, Kv, Kv_ml,
                      Kv_extra_bbl, Hmix, Hbbl, &tb);
        if (m > 0.0) tauy_bot[c2] = tb;
    }
}

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
    const size_t lds = 2 * (size_t)(d.nk + 1) * VV_BLOCK * sizeof(double);
    if ((long long)lds <= cs->max_lds) {
        dim3 b(VV_BLOCK), g((ni + VV_BLOCK - 1) / VV_BLOCK, rows);
        const bool keep = cs->keep_coefs != 0;
        k_vert_visc_lds<<<g, b, lds, (hipStream_t)stream>>>(
            d, u, v, h, dt, taux, tauy, ray_u_bot, ray_v_bot, cs->mask2dCu, cs->mask2dCv, keep ? cs->a_u : nullptr,
            keep ? cs->a_v : nullptr, keep ? cs->h_u : nullptr, keep ? cs->h_v : nullptr, cs->taux_bot,
            cs->tauy_bot, cs->Kv, cs->Kv_ml, cs->Kv_extra_bbl, cs->Hmix, cs->Hbbl);
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
