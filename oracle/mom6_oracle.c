/*
 * mom6_oracle.c -- CPU restatement of the MOM6 mini-app serial split-RK2 step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see mom6_oracle.h.  Every routine
 * cites the reference subroutine it restates.  Arithmetic follows Fortran's
 * left-to-right association and explicit parentheses; default-kind (REAL(4))
 * literals and locals of the serial sources are reproduced through float.
 * Build with -ffp-contract=off and without fast-math.
 */
#include "mom6_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Physical constants, src/common/mom6_types.F90:80-84 */
#define G_EARTH 9.80
#define RHO_0 1035.0
#define OMEGA 7.2921e-5
#define PI_REF 3.14159265358979323846
#define EARTH_RADIUS 6.371e6

#define ISD 1
#define JSD 1

/* Fortran-indexed accessors (1-based, column-major). */
#define A2(a, i, j) ((a)[(size_t)((j) - 1) * m->n1 + (size_t)((i) - 1)])
#define A3(a, i, j, k) ((a)[(((size_t)(k) - 1) * m->n2 + (size_t)((j) - 1)) * m->n1 + (size_t)((i) - 1)])
#define V1(a, i) ((a)[(i) - 1])                       /* per-row vector a(isd:ied) */
#define R2(a, i, k) ((a)[((size_t)(k) - 1) * m->n1 + (size_t)((i) - 1)])  /* a(isd:ied, nz) */

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
typedef struct named { const char *name; double **p; long n; } named;

struct om_model {
    om_config c;
    double *diag_work2d, *diag_ke, *diag_mass;   /* diagnostics: work_2d, KE, mass */
    double diag_posted[8];                       /* the last diffu_sum, diffv_sum statistics */
    int ni, nj, nk, n1, n2;
    int isd, ied, jsd, jed, isc, iec, jsc, jec;
    size_t N2, N3;
    double dx;
    int first_direction;

    /* grid (mom6_types.F90:26-57) */
    double *IareaT, *areaT, *dxT, *dyT, *IdxT, *IdyT, *dxCu, *dyCu, *dy_Cu, *IdxCu, *IdyCu;
    double *dxCv, *dyCv, *IdxCv, *IdyCv, *IareaBu, *areaBu, *CoriolisBu, *IareaCu, *IareaCv;
    double *dxBu, *dyBu, *IdxBu, *IdyBu, *bathyT, *mask2dT, *mask2dBu, *mask2dCu, *mask2dCv;

    /* driver state (rk2_serial_driver.F90:59-76) */
    double *u, *v, *h, *h0, *htmp, *uh, *vh, *CAu, *CAv, *up, *vp, *diffu, *diffv;
    double *eta, *ubt, *vbt, *ubt_av, *vbt_av, *eta_av;
    double *uhbt, *u_cor, *du_cor, *por_face_areaU, *visc_rem_u;  /* continuity_init */
    double *vhbt;            /* the meridional pass's barotropic-transport target */
    double *taux, *tauy, *Ray_u, *Ray_v;                             /* forces / visc */
    /* BT_cont (mom6_types.F90:89-110) */
    double *FA_u_EE, *FA_u_E0, *FA_u_W0, *FA_u_WW, *uBT_WW, *uBT_EE, *BT_h_u;

    /* continuity_CS (mom6_continuity_serial.F90:17-48) */
    int upwind_1st, monotonic, simple_2nd, aggress_adjust, vol_CFL, better_iter;
    int use_visc_rem_max, marginal_faces;
    double tol_eta, tol_vel, CFL_limit_adjust;
    double *cont_hW, *cont_hE;   /* automatic arrays h_W/h_E of continuity_PPM */

    /* coriolis_CS (mom6_coriolis_serial.F90:24-40) */
    double *cor_dvdx, *cor_dudy, *cor_rel_vort, *cor_abs_vort, *cor_q, *cor_Ih_q;
    double *cor_hArea_u, *cor_hArea_v, *cor_Area_q, *cor_KE, *cor_a, *cor_b, *cor_c, *cor_d;

    /* barotropic_CS (mom6_barotropic_serial.F90:20-58) */
    double bt_dtbt, bt_bebt, bt_dgeo_de;
    int bt_nstep;
    double *bt_eta, *bt_eta_pred, *bt_ubt, *bt_vbt, *bt_ubt_prev, *bt_vbt_prev, *bt_uhbt, *bt_vhbt;
    double *bt_PFu, *bt_PFv, *bt_Cor_u, *bt_Cor_v, *bt_Datu, *bt_Datv;
    double *bt_gtot_E, *bt_gtot_W, *bt_gtot_N, *bt_gtot_S, *bt_f_4_u, *bt_f_4_v;
    double *bt_rem_u, *bt_rem_v, *bt_ubt_av, *bt_vbt_av;

    /* vert_visc_CS (mom6_vert_visc_serial.F90:27-55) */
    double vv_Kv, vv_Kv_extra_bbl, vv_Hmix, vv_Hbbl, vv_Kv_ml;
    double *vv_a_u, *vv_a_v, *vv_h_u, *vv_h_v, *vv_taux_bot, *vv_tauy_bot;

    /* hor_visc_CS (mom6_hor_visc_serial.F90:32-152), the subset the path uses */
    int hv_Laplacian, hv_biharmonic, hv_Smag_Kh, hv_Smag_Ah, hv_no_slip;
    int hv_bound_Kh, hv_better_bound_Kh, hv_bound_Ah, hv_better_bound_Ah, hv_use_land_mask;
    double hv_Kh_bg_min, hv_h_neglect;
    double *hv_Kh_bg_xx, *hv_Kh_bg_xy, *hv_Ah_bg_xx, *hv_Ah_bg_xy;
    double *hv_Kh_Max_xx, *hv_Kh_Max_xy, *hv_Ah_Max_xx, *hv_Ah_Max_xy;
    double *hv_Laplac2_const_xx, *hv_Laplac2_const_xy, *hv_Biharm_const_xx, *hv_Biharm_const_xy;
    double *hv_dx2h, *hv_dy2h, *hv_dx2q, *hv_dy2q, *hv_DX_dyT, *hv_DY_dxT, *hv_DX_dyBu, *hv_DY_dxBu;
    double *hv_Idx2dyCu, *hv_Idxdy2u, *hv_Idx2dyCv, *hv_Idxdy2v, *hv_reduction_xx, *hv_reduction_xy;
    double *hv_dudx, *hv_dvdy, *hv_dvdx, *hv_dudy, *hv_sh_xx, *hv_sh_xy, *hv_str_xx, *hv_str_xy;
    double *hv_bhstr_xx, *hv_bhstr_xy, *hv_h_u, *hv_h_v, *hv_hq, *hv_Kh, *hv_Ah, *hv_Shear_mag;
    double *hv_Del2u, *hv_Del2v, *hv_hrat_min, *hv_visc_bound_rem;
    int hv_Leith_Kh, hv_Leith_Ah;
    double *hv_Laplac3_const_xx, *hv_Laplac3_const_xy, *hv_Biharm6_const_xx, *hv_Biharm6_const_xy;
    double *hv_vort_xy, *hv_vort_xy_dx, *hv_vort_xy_dy, *hv_grad_vort_mag_h, *hv_grad_vort_mag_q;
    double *hv_vert_vort_mag, *hv_Del2vort_q;

    /* meridional pass: the transposed model it runs the zonal pass on, and the
     * meridional outputs (v_cor, dv_cor, BT_cont v-side) */
    struct om_model *mt;
    double *v_cor, *dv_cor, *FA_v_NN, *FA_v_N0, *FA_v_S0, *FA_v_SS, *vBT_SS, *vBT_NN, *BT_h_v;

    named *tab;
    int ntab, captab;
};

/* ------------------------------------------------------------------------ */
static double *reg(om_model *m, const char *name, double **slot, size_t n, double fill)
{
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (fill != 0.0)
        for (size_t q = 0; q < n; ++q) p[q] = fill;
    *slot = p;
    if (m->ntab == m->captab) {
        m->captab = m->captab ? 2 * m->captab : 256;
        m->tab = (named *)realloc(m->tab, (size_t)m->captab * sizeof(named));
    }
    m->tab[m->ntab].name = name;
    m->tab[m->ntab].p = slot;
    m->tab[m->ntab].n = (long)n;
    m->ntab++;
    return p;
}

void om_config_default(om_config *c)
{
    memset(c, 0, sizeof(*c));
    c->ni = 180; c->nj = 180; c->nk = 75; c->bt_nsteps = 30;
    c->dt = 300.0; c->dx_km = 10.0;                       /* rk2_serial_driver.F90:91-93,135 */
    c->bc_x = OM_BC_STATIC; c->bc_y = OM_BC_STATIC;
    c->coriolis_mode = 0; c->f0 = 1.0e-4; c->beta = 0.0;
    c->Kh = 100.0; c->Ah = 1.0e10;                        /* rk2_serial_driver.F90:143, hor_visc:197-200 */
    c->laplacian = 1; c->biharmonic = 0;
    c->Kv = 1.0e-4; c->Kv_ml = 1.0e-2; c->Kv_extra_bbl = 1.0e-2; c->Hmix = 50.0; c->Hbbl = 10.0;
    c->ray_bottom = 1.0e-4;                               /* rk2_serial_driver.F90:536-537 */
    c->coriolis_scheme = 1;
    c->monotonic = 0;                                     /* never set by continuity_init */
}

/* ---------------------------------------------------------------- grid --- */
/* init_ocean_grid, mom6_types.F90:130-258 */
static void grid_init(om_model *m)
{
    const double dx_m = m->c.dx_km * 1000.0;
    double f0, beta;
    m->dx = dx_m;
    if (m->c.coriolis_mode == 0) {
        f0 = 2.0 * OMEGA * sin(45.0 * PI_REF / 180.0);
        beta = 2.0 * OMEGA * cos(45.0 * PI_REF / 180.0) / EARTH_RADIUS;
    } else {
        f0 = m->c.f0;
        beta = m->c.beta;
    }
    for (int j = m->jsd; j <= m->jed; ++j)
        for (int i = m->isd; i <= m->ied; ++i) {
            A2(m->areaT, i, j) = dx_m * dx_m;
            A2(m->IareaT, i, j) = 1.0 / (dx_m * dx_m);
            A2(m->dxT, i, j) = dx_m;
            A2(m->dyT, i, j) = dx_m;
            A2(m->IdxT, i, j) = 1.0 / dx_m;
            A2(m->IdyT, i, j) = 1.0 / dx_m;
            A2(m->dxCu, i, j) = dx_m;
            A2(m->dyCu, i, j) = dx_m;
            A2(m->dy_Cu, i, j) = dx_m;
            A2(m->IdxCu, i, j) = 1.0 / dx_m;
            A2(m->IdyCu, i, j) = 1.0 / dx_m;
            A2(m->dxCv, i, j) = dx_m;
            A2(m->dyCv, i, j) = dx_m;
            A2(m->IdxCv, i, j) = 1.0 / dx_m;
            A2(m->IdyCv, i, j) = 1.0 / dx_m;
            A2(m->areaBu, i, j) = dx_m * dx_m;
            A2(m->IareaBu, i, j) = 1.0 / (dx_m * dx_m);
            A2(m->IareaCu, i, j) = 1.0 / (dx_m * dx_m);
            A2(m->IareaCv, i, j) = 1.0 / (dx_m * dx_m);
            A2(m->dxBu, i, j) = dx_m;
            A2(m->dyBu, i, j) = dx_m;
            A2(m->IdxBu, i, j) = 1.0 / dx_m;
            A2(m->IdyBu, i, j) = 1.0 / dx_m;
            A2(m->mask2dT, i, j) = 1.0;
            A2(m->mask2dBu, i, j) = 1.0;
            A2(m->mask2dCu, i, j) = 1.0;
            A2(m->mask2dCv, i, j) = 1.0;
            A2(m->bathyT, i, j) = 4000.0;
            A2(m->CoriolisBu, i, j) = f0 + beta * ((double)(j - 1 - m->nj / 2) - 0.5) * dx_m;
        }
    if (m->c.bc_y == OM_BC_WALL) {
        /* Land rows outside the compute rows; v and q points on the two wall
         * lines (J = jsc-1 and J = jec) are closed. */
        for (int j = m->jsd; j <= m->jed; ++j)
            for (int i = m->isd; i <= m->ied; ++i) {
                if (j < m->jsc || j > m->jec) {
                    A2(m->mask2dT, i, j) = 0.0;
                    A2(m->mask2dCu, i, j) = 0.0;
                }
                if (j < m->jsc || j >= m->jec) {
                    A2(m->mask2dCv, i, j) = 0.0;
                    A2(m->mask2dBu, i, j) = 0.0;
                }
            }
    }
}

/* ------------------------------------------------------------ halo fill --- */
/* Periodic refill of the 3-wide halo (the single-domain counterpart of
 * halo_exchange_3d/2d in src/common/mom6_mpi_halo.F90 on a periodic grid).
 * x first over all rows, then y over all columns (corners follow). */
static void halo_fill(om_model *m, double *a, int nk)
{
    const int P = m->iec - m->isc + 1, Q = m->jec - m->jsc + 1;
    for (int k = 1; k <= nk; ++k) {
        if (m->c.bc_x == OM_BC_PERIODIC)
            for (int j = m->jsd; j <= m->jed; ++j) {
                for (int i = m->isd; i < m->isc; ++i) A3(a, i, j, k) = A3(a, i + P, j, k);
                for (int i = m->iec + 1; i <= m->ied; ++i) A3(a, i, j, k) = A3(a, i - P, j, k);
            }
        if (m->c.bc_y == OM_BC_PERIODIC)
            for (int i = m->isd; i <= m->ied; ++i) {
                for (int j = m->jsd; j < m->jsc; ++j) A3(a, i, j, k) = A3(a, i, j + Q, k);
                for (int j = m->jec + 1; j <= m->jed; ++j) A3(a, i, j, k) = A3(a, i, j - Q, k);
            }
    }
}
void om_halo_fill(om_model *m, double *a, int nk) { halo_fill(m, a, nk); }

static int periodic(const om_model *m)
{
    return m->c.bc_x == OM_BC_PERIODIC || m->c.bc_y == OM_BC_PERIODIC;
}

/* ---------------------------------------------------------- continuity --- */
/* PPM_limit_pos, mom6_continuity_serial.F90:1253-1294 (REAL(4) curv, dh, scale) */
static void ppm_limit_pos(om_model *m, const double *h_in, double *h_L, double *h_R, double h_min,
                          int iis, int iie, int jis, int jie, int k)
{
    for (int j = jis; j <= jie; ++j)
        for (int i = iis; i <= iie; ++i) {
            const double hi = A3(h_in, i, j, k);
            double *hl = &A3(h_L, i, j, k), *hr = &A3(h_R, i, j, k);
            float curv = (float)(3.0 * ((*hl + *hr) - 2.0 * hi));
            if ((double)curv > 0.0) {
                float dh = (float)(*hr - *hl);
                if (fabsf(dh) < curv) {
                    if (hi <= h_min) {
                        *hl = hi;
                        *hr = hi;
                    } else {
                        const double lhs = (double)(12.0f * curv) * (hi - h_min);
                        const float rhs = curv * curv + 3.0f * (dh * dh);
                        if (lhs < (double)rhs) {
                            float scale = (float)(lhs / (double)rhs);
                            *hl = hi + (double)scale * (*hl - hi);
                            *hr = hi + (double)scale * (*hr - hi);
                        }
                    }
                }
            }
        }
}

/* PPM_limit_CW84, mom6_continuity_serial.F90:1298-1338 (REAL(4) locals) */
static void ppm_limit_cw84(om_model *m, const double *h_in, double *h_L, double *h_R,
                           int iis, int iie, int jis, int jie, int k)
{
    for (int j = jis; j <= jie; ++j)
        for (int i = iis; i <= iie; ++i) {
            double *hl = &A3(h_L, i, j, k), *hr = &A3(h_R, i, j, k);
            const float h_i = (float)A3(h_in, i, j, k);
            if ((*hr - (double)h_i) * ((double)h_i - *hl) <= 0.0) {
                *hl = (double)h_i;
                *hr = (double)h_i;
            } else {
                const float RLdiff = (float)(*hr - *hl);
                const float RLmean = (float)(0.5 * (*hr + *hl));
                const float FunFac = 6.0f * RLdiff * (h_i - RLmean);
                const float RLdiff2 = RLdiff * RLdiff;
                if (FunFac > RLdiff2) *hl = (double)(3.0f * h_i) - 2.0 * *hr;
                if (FunFac < -RLdiff2) *hr = (double)(3.0f * h_i) - 2.0 * *hl;
            }
        }
}

/* PPM_reconstruction_x, mom6_continuity_serial.F90:1124-1247 (one layer k) */
static void ppm_reconstruction_x(om_model *m, const double *h_in, double *h_W, double *h_E,
                                 double h_min, int k, double *slp)
{
    const int isl = m->isc - 1, iel = m->iec + 1, jsl = m->jsc, jel = m->jec;
    const double oneSixth = 1.0 / 6.0;
    if (m->simple_2nd) {
        for (int j = jsl; j <= jel; ++j)
            for (int i = isl; i <= iel; ++i) {
                const double h_im1 = A2(m->mask2dT, i - 1, j) * A3(h_in, i - 1, j, k) +
                                     (1.0 - A2(m->mask2dT, i - 1, j)) * A3(h_in, i, j, k);
                const double h_ip1 = A2(m->mask2dT, i + 1, j) * A3(h_in, i + 1, j, k) +
                                     (1.0 - A2(m->mask2dT, i + 1, j)) * A3(h_in, i, j, k);
                A3(h_W, i, j, k) = 0.5 * (h_im1 + A3(h_in, i, j, k));
                A3(h_E, i, j, k) = 0.5 * (h_ip1 + A3(h_in, i, j, k));
            }
    } else {
        for (int j = jsl; j <= jel; ++j)
            for (int i = isl - 1; i <= iel + 1; ++i) {
                double s;
                if ((A2(m->mask2dT, i - 1, j) * A2(m->mask2dT, i, j) * A2(m->mask2dT, i + 1, j)) == 0.0) {
                    s = 0.0;
                } else {
                    const double hp = A3(h_in, i + 1, j, k), hm = A3(h_in, i - 1, j, k), h0 = A3(h_in, i, j, k);
                    s = 0.5 * (hp - hm);
                    const double dMx = dmax(dmax(hp, hm), h0) - h0;
                    const double dMn = h0 - dmin(dmin(hp, hm), h0);
                    s = copysign(1.0, s) * dmin(fabs(s), 2.0 * dmin(dMx, dMn));
                }
                A2(slp, i, j) = s;
            }
        for (int j = jsl; j <= jel; ++j)
            for (int i = isl; i <= iel; ++i) {
                const double h_im1 = A2(m->mask2dT, i - 1, j) * A3(h_in, i - 1, j, k) +
                                     (1.0 - A2(m->mask2dT, i - 1, j)) * A3(h_in, i, j, k);
                const double h_ip1 = A2(m->mask2dT, i + 1, j) * A3(h_in, i + 1, j, k) +
                                     (1.0 - A2(m->mask2dT, i + 1, j)) * A3(h_in, i, j, k);
                A3(h_W, i, j, k) = 0.5 * (h_im1 + A3(h_in, i, j, k)) + oneSixth * (A2(slp, i - 1, j) - A2(slp, i, j));
                A3(h_E, i, j, k) = 0.5 * (h_ip1 + A3(h_in, i, j, k)) + oneSixth * (A2(slp, i, j) - A2(slp, i + 1, j));
            }
    }
    if (m->monotonic)
        ppm_limit_cw84(m, h_in, h_W, h_E, isl, iel, jsl, jel, k);
    else
        ppm_limit_pos(m, h_in, h_W, h_E, h_min, isl, iel, jsl, jel, k);
}

/* zonal_edge_thickness, mom6_continuity_serial.F90:179-215 */
static void zonal_edge_thickness(om_model *m, const double *h_in, double *h_W, double *h_E)
{
    if (m->upwind_1st) {
        for (int k = 1; k <= m->nk; ++k)
            for (int j = m->jsc; j <= m->jec; ++j)
                for (int i = m->isc - 1; i <= m->iec + 1; ++i) {
                    A3(h_W, i, j, k) = A3(h_in, i, j, k);
                    A3(h_E, i, j, k) = A3(h_in, i, j, k);
                }
    } else {
        double *slp = (double *)calloc(m->N2, sizeof(double));
        for (int k = 1; k <= m->nk; ++k)
            ppm_reconstruction_x(m, h_in, h_W, h_E, 2.0 * 1.0e-10, k, slp);
        free(slp);
    }
}

static double ratio_max(double a, double b, double maxrat)  /* :1341-1352 */
{
    return (fabs(a) > fabs(maxrat * b)) ? maxrat : a / b;
}

/* zonal_flux_layer, mom6_continuity_serial.F90:516-598.  Row vectors are
 * indexed I = isd..ied; u/h/hW/hE/visc_rem/por are the j-row (layer k). */
static void zonal_flux_layer(om_model *m, const double *u, const double *h, const double *h_W,
                             const double *h_E, double *uh, double *duhdu, const double *visc_rem,
                             double dt, int j, int ish, int ieh, const int *do_I, const double *por)
{
    for (int I = ish - 1; I <= ieh; ++I) {
        if (!V1(do_I, I)) continue;
        const int i = I;
        double CFL, curv_3, h_marg;
        const double uI = V1(u, I);
        if (uI > 0.0) {
            if (m->vol_CFL) CFL = (uI * dt) * (A2(m->dy_Cu, I, j) * A2(m->IareaT, i, j));
            else CFL = uI * dt * A2(m->IdxT, i, j);
            curv_3 = (V1(h_W, i) + V1(h_E, i)) - 2.0 * V1(h, i);
            V1(uh, I) = (A2(m->dy_Cu, I, j) * V1(por, I)) * uI *
                        (V1(h_E, i) + CFL * (0.5 * (V1(h_W, i) - V1(h_E, i)) + curv_3 * (CFL - 1.5)));
            h_marg = V1(h_E, i) + CFL * ((V1(h_W, i) - V1(h_E, i)) + 3.0 * curv_3 * (CFL - 1.0));
        } else if (uI < 0.0) {
            if (m->vol_CFL) CFL = (-uI * dt) * (A2(m->dy_Cu, I, j) * A2(m->IareaT, i + 1, j));
            else CFL = -(uI * dt * A2(m->IdxT, i + 1, j));
            curv_3 = (V1(h_W, i + 1) + V1(h_E, i + 1)) - 2.0 * V1(h, i + 1);
            V1(uh, I) = (A2(m->dy_Cu, I, j) * V1(por, I)) * uI *
                        (V1(h_W, i + 1) + CFL * (0.5 * (V1(h_E, i + 1) - V1(h_W, i + 1)) + curv_3 * (CFL - 1.5)));
            h_marg = V1(h_W, i + 1) + CFL * ((V1(h_E, i + 1) - V1(h_W, i + 1)) + 3.0 * curv_3 * (CFL - 1.0));
        } else {
            V1(uh, I) = 0.0;
            h_marg = 0.5 * (V1(h_W, i + 1) + V1(h_E, i));
        }
        V1(duhdu, I) = (A2(m->dy_Cu, I, j) * V1(por, I)) * h_marg * V1(visc_rem, I);
    }
}

/* Row j, layer k slice pointers of a 3D array (the Fortran a(:, j, k) section). */
#define ROW(a, j, k) (&A3(a, 1, j, k))

/* zonal_flux_adjust, mom6_continuity_serial.F90:739-916.  uh_3d may be NULL. */
static void zonal_flux_adjust(om_model *m, const double *u, const double *h_in, const double *h_W,
                              const double *h_E, const double *uhbt, const double *uh_tot_0,
                              const double *duhdu_tot_0, double *du, const double *du_max_CFL,
                              const double *du_min_CFL, double dt, const double *visc_rem, int j,
                              int ish, int ieh, const int *do_I_in, const double *por, double *uh_3d)
{
    const int n1 = m->n1, nz = m->nk;
    double *uh_aux = (double *)calloc((size_t)n1 * nz, sizeof(double));
    double *duhdu = (double *)calloc((size_t)n1 * nz, sizeof(double));
    double *uh_err = (double *)calloc(n1, sizeof(double));
    double *uh_err_best = (double *)calloc(n1, sizeof(double));
    double *u_new = (double *)calloc(n1, sizeof(double));
    double *duhdu_tot = (double *)calloc(n1, sizeof(double));
    double *du_min = (double *)calloc(n1, sizeof(double));
    double *du_max = (double *)calloc(n1, sizeof(double));
    int *do_I = (int *)calloc(n1, sizeof(int));
    const int max_itts = 20;

    if (uh_3d)
        for (int k = 1; k <= nz; ++k)
            for (int I = ish - 1; I <= ieh; ++I) R2(uh_aux, I, k) = A3(uh_3d, I, j, k);

    for (int I = ish - 1; I <= ieh; ++I) {
        V1(du, I) = 0.0;
        V1(do_I, I) = V1(do_I_in, I);
        V1(du_max, I) = V1(du_max_CFL, I);
        V1(du_min, I) = V1(du_min_CFL, I);
        V1(uh_err, I) = V1(uh_tot_0, I) - V1(uhbt, I);
        V1(duhdu_tot, I) = V1(duhdu_tot_0, I);
        V1(uh_err_best, I) = fabs(V1(uh_err, I));
    }

    for (int itt = 1; itt <= max_itts; ++itt) {
        double tol_eta;
        /* default-real literals 1e-6, 1e-4, 1e-2 (:826-830) */
        if (itt <= 1) tol_eta = (double)1e-6f * m->tol_eta;
        else if (itt == 2) tol_eta = (double)1e-4f * m->tol_eta;
        else if (itt == 3) tol_eta = (double)1e-2f * m->tol_eta;
        else tol_eta = m->tol_eta;
        const double tol_vel = m->tol_vel;

        for (int I = ish - 1; I <= ieh; ++I) {
            if (V1(uh_err, I) > 0.0) V1(du_max, I) = V1(du, I);
            else if (V1(uh_err, I) < 0.0) V1(du_min, I) = V1(du, I);
            else V1(do_I, I) = 0;
        }
        int domore = 0;
        for (int I = ish - 1; I <= ieh; ++I) {
            if (!V1(do_I, I)) continue;
            const int i = I;
            const double aerr = fabs(V1(uh_err, I));
            if ((dt * dmin(A2(m->IareaT, i, j), A2(m->IareaT, i + 1, j)) * aerr > tol_eta) ||
                (m->better_iter && ((aerr > tol_vel * V1(duhdu_tot, I)) || (aerr > V1(uh_err_best, I))))) {
                const double ddu = -V1(uh_err, I) / V1(duhdu_tot, I);
                const double du_prev = V1(du, I);
                V1(du, I) = V1(du, I) + ddu;
                if (fabs(ddu) < (double)1.0e-15f * fabs(V1(du, I))) {
                    V1(do_I, I) = 0;
                } else if (ddu > 0.0) {
                    if (V1(du, I) >= V1(du_max, I)) {
                        V1(du, I) = 0.5 * (du_prev + V1(du_max, I));
                        if (V1(du_max, I) - du_prev < (double)1.0e-15f * fabs(V1(du, I))) V1(do_I, I) = 0;
                    }
                } else {
                    if (V1(du, I) <= V1(du_min, I)) {
                        V1(du, I) = 0.5 * (du_prev + V1(du_min, I));
                        if (du_prev - V1(du_min, I) < (double)1.0e-15f * fabs(V1(du, I))) V1(do_I, I) = 0;
                    }
                }
                if (V1(do_I, I)) domore = 1;
            } else {
                V1(do_I, I) = 0;
            }
        }
        if (!domore) break;

        if ((itt < max_itts) || uh_3d) {
            for (int k = 1; k <= nz; ++k) {
                for (int I = ish - 1; I <= ieh; ++I)
                    V1(u_new, I) = A3(u, I, j, k) + V1(du, I) * R2(visc_rem, I, k);
                zonal_flux_layer(m, u_new, ROW(h_in, j, k), ROW(h_W, j, k), ROW(h_E, j, k),
                                 &R2(uh_aux, 1, k), &R2(duhdu, 1, k), &R2(visc_rem, 1, k), dt, j,
                                 ish, ieh, do_I, ROW(por, j, k));
            }
        }
        if (itt < max_itts) {
            for (int I = ish - 1; I <= ieh; ++I) {
                V1(uh_err, I) = -V1(uhbt, I);
                V1(duhdu_tot, I) = 0.0;
            }
            for (int k = 1; k <= nz; ++k)
                for (int I = ish - 1; I <= ieh; ++I) {
                    V1(uh_err, I) = V1(uh_err, I) + R2(uh_aux, I, k);
                    V1(duhdu_tot, I) = V1(duhdu_tot, I) + R2(duhdu, I, k);
                }
            for (int I = ish - 1; I <= ieh; ++I)
                V1(uh_err_best, I) = dmin(V1(uh_err_best, I), fabs(V1(uh_err, I)));
        }
    }
    if (uh_3d)
        for (int k = 1; k <= nz; ++k)
            for (int I = ish - 1; I <= ieh; ++I) A3(uh_3d, I, j, k) = R2(uh_aux, I, k);

    free(uh_aux); free(duhdu); free(uh_err); free(uh_err_best); free(u_new);
    free(duhdu_tot); free(du_min); free(du_max); free(do_I);
}

/* set_zonal_BT_cont, mom6_continuity_serial.F90:920-1121 */
static void set_zonal_BT_cont(om_model *m, const double *u, const double *h_in, const double *h_W,
                              const double *h_E, const double *uh_tot_0, const double *duhdu_tot_0,
                              const double *du_max_CFL, const double *du_min_CFL, double dt,
                              const double *visc_rem, const double *visc_rem_max, int j, int ish,
                              int ieh, const int *do_I, const double *por)
{
    const int n1 = m->n1, nz = m->nk;
    double *du0 = calloc(n1, sizeof(double)), *duL = calloc(n1, sizeof(double));
    double *duR = calloc(n1, sizeof(double)), *zeros = calloc(n1, sizeof(double));
    double *du_CFL = calloc(n1, sizeof(double)), *u_L = calloc(n1, sizeof(double));
    double *u_R = calloc(n1, sizeof(double)), *u_0 = calloc(n1, sizeof(double));
    double *duhdu_L = calloc(n1, sizeof(double)), *duhdu_R = calloc(n1, sizeof(double));
    double *duhdu_0 = calloc(n1, sizeof(double)), *uh_L = calloc(n1, sizeof(double));
    double *uh_R = calloc(n1, sizeof(double)), *uh_0 = calloc(n1, sizeof(double));
    double *FAmt_L = calloc(n1, sizeof(double)), *FAmt_R = calloc(n1, sizeof(double));
    double *FAmt_0 = calloc(n1, sizeof(double)), *uhtot_L = calloc(n1, sizeof(double));
    double *uhtot_R = calloc(n1, sizeof(double));
    const double Idt = 1.0 / dt;
    const double min_visc_rem = (double)0.1f;   /* default-real literals (:993-994) */
    const double CFL_min = (double)1e-6f;

    for (int I = ish - 1; I <= ieh; ++I) V1(zeros, I) = 0.0;
    zonal_flux_adjust(m, u, h_in, h_W, h_E, zeros, uh_tot_0, duhdu_tot_0, du0, du_max_CFL,
                      du_min_CFL, dt, visc_rem, j, ish, ieh, do_I, por, NULL);

    int domore = 0;
    for (int I = ish - 1; I <= ieh; ++I) {
        if (V1(do_I, I)) domore = 1;
        V1(du_CFL, I) = (CFL_min * Idt) * A2(m->dxCu, I, j);
        V1(duR, I) = dmin(0.0, V1(du0, I) - V1(du_CFL, I));
        V1(duL, I) = dmax(0.0, V1(du0, I) + V1(du_CFL, I));
        V1(FAmt_L, I) = 0.0; V1(FAmt_R, I) = 0.0; V1(FAmt_0, I) = 0.0;
        V1(uhtot_L, I) = 0.0; V1(uhtot_R, I) = 0.0;
    }
    if (!domore) {
        for (int I = ish - 1; I <= ieh; ++I) {
            A2(m->FA_u_W0, I, j) = 0.0; A2(m->FA_u_WW, I, j) = 0.0;
            A2(m->FA_u_E0, I, j) = 0.0; A2(m->FA_u_EE, I, j) = 0.0;
            A2(m->uBT_WW, I, j) = 0.0; A2(m->uBT_EE, I, j) = 0.0;
        }
        goto done;
    }
    for (int k = 1; k <= nz; ++k)
        for (int I = ish - 1; I <= ieh; ++I) {
            if (!V1(do_I, I)) continue;
            const double vr = R2(visc_rem, I, k);
            const double visc_rem_lim = dmax(vr, min_visc_rem * V1(visc_rem_max, I));
            if (visc_rem_lim > 0.0) {
                const double uu = A3(u, I, j, k);
                if (uu + V1(duR, I) * visc_rem_lim > -V1(du_CFL, I) * vr)
                    V1(duR, I) = -(uu + V1(du_CFL, I) * vr) / visc_rem_lim;
                if (uu + V1(duL, I) * visc_rem_lim < V1(du_CFL, I) * vr)
                    V1(duL, I) = -(uu - V1(du_CFL, I) * vr) / visc_rem_lim;
            }
        }
    for (int k = 1; k <= nz; ++k) {
        for (int I = ish - 1; I <= ieh; ++I)
            if (V1(do_I, I)) {
                V1(u_L, I) = A3(u, I, j, k) + V1(duL, I) * R2(visc_rem, I, k);
                V1(u_R, I) = A3(u, I, j, k) + V1(duR, I) * R2(visc_rem, I, k);
                V1(u_0, I) = A3(u, I, j, k) + V1(du0, I) * R2(visc_rem, I, k);
            }
        zonal_flux_layer(m, u_0, ROW(h_in, j, k), ROW(h_W, j, k), ROW(h_E, j, k), uh_0, duhdu_0,
                         &R2(visc_rem, 1, k), dt, j, ish, ieh, do_I, ROW(por, j, k));
        zonal_flux_layer(m, u_L, ROW(h_in, j, k), ROW(h_W, j, k), ROW(h_E, j, k), uh_L, duhdu_L,
                         &R2(visc_rem, 1, k), dt, j, ish, ieh, do_I, ROW(por, j, k));
        zonal_flux_layer(m, u_R, ROW(h_in, j, k), ROW(h_W, j, k), ROW(h_E, j, k), uh_R, duhdu_R,
                         &R2(visc_rem, 1, k), dt, j, ish, ieh, do_I, ROW(por, j, k));
        for (int I = ish - 1; I <= ieh; ++I)
            if (V1(do_I, I)) {
                V1(FAmt_0, I) = V1(FAmt_0, I) + V1(duhdu_0, I);
                V1(FAmt_L, I) = V1(FAmt_L, I) + V1(duhdu_L, I);
                V1(FAmt_R, I) = V1(FAmt_R, I) + V1(duhdu_R, I);
                V1(uhtot_L, I) = V1(uhtot_L, I) + V1(uh_L, I);
                V1(uhtot_R, I) = V1(uhtot_R, I) + V1(uh_R, I);
            }
    }
    for (int I = ish - 1; I <= ieh; ++I) {
        if (V1(do_I, I)) {
            double FA_0 = V1(FAmt_0, I), FA_avg = V1(FAmt_0, I);
            if ((V1(duL, I) - V1(du0, I)) != 0.0) FA_avg = V1(uhtot_L, I) / (V1(duL, I) - V1(du0, I));
            if (FA_avg > dmax(FA_0, V1(FAmt_L, I))) FA_avg = dmax(FA_0, V1(FAmt_L, I));
            else if (FA_avg < dmin(FA_0, V1(FAmt_L, I))) FA_0 = FA_avg;
            A2(m->FA_u_W0, I, j) = FA_0;
            A2(m->FA_u_WW, I, j) = V1(FAmt_L, I);
            if (fabs(FA_0 - V1(FAmt_L, I)) <= (double)1e-12f * FA_0) A2(m->uBT_WW, I, j) = 0.0;
            else A2(m->uBT_WW, I, j) = (1.5 * (V1(duL, I) - V1(du0, I))) *
                                       ((V1(FAmt_L, I) - FA_avg) / (V1(FAmt_L, I) - FA_0));

            FA_0 = V1(FAmt_0, I); FA_avg = V1(FAmt_0, I);
            if ((V1(duR, I) - V1(du0, I)) != 0.0) FA_avg = V1(uhtot_R, I) / (V1(duR, I) - V1(du0, I));
            if (FA_avg > dmax(FA_0, V1(FAmt_R, I))) FA_avg = dmax(FA_0, V1(FAmt_R, I));
            else if (FA_avg < dmin(FA_0, V1(FAmt_R, I))) FA_0 = FA_avg;
            A2(m->FA_u_E0, I, j) = FA_0;
            A2(m->FA_u_EE, I, j) = V1(FAmt_R, I);
            if (fabs(V1(FAmt_R, I) - FA_0) <= (double)1e-12f * FA_0) A2(m->uBT_EE, I, j) = 0.0;
            else A2(m->uBT_EE, I, j) = (1.5 * (V1(duR, I) - V1(du0, I))) *
                                       ((V1(FAmt_R, I) - FA_avg) / (V1(FAmt_R, I) - FA_0));
        } else {
            A2(m->FA_u_W0, I, j) = 0.0; A2(m->FA_u_WW, I, j) = 0.0;
            A2(m->FA_u_E0, I, j) = 0.0; A2(m->FA_u_EE, I, j) = 0.0;
            A2(m->uBT_WW, I, j) = 0.0; A2(m->uBT_EE, I, j) = 0.0;
        }
    }
done:
    free(du0); free(duL); free(duR); free(zeros); free(du_CFL); free(u_L); free(u_R); free(u_0);
    free(duhdu_L); free(duhdu_R); free(duhdu_0); free(uh_L); free(uh_R); free(uh_0);
    free(FAmt_L); free(FAmt_R); free(FAmt_0); free(uhtot_L); free(uhtot_R);
}

/* zonal_flux_thickness, mom6_continuity_serial.F90:602-735 */
static void zonal_flux_thickness(om_model *m, const double *u, const double *h, const double *h_W,
                                 const double *h_E, double *h_u, double dt)
{
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int I = m->isc - 1; I <= m->iec; ++I) {
                const int i = I;
                const double uu = A3(u, I, j, k);
                double CFL, curv_3, h_avg, h_marg;
                if (uu > 0.0) {
                    if (m->vol_CFL) CFL = (uu * dt) * (A2(m->dy_Cu, I, j) * A2(m->IareaT, i, j));
                    else CFL = uu * dt * A2(m->IdxT, i, j);
                    curv_3 = (A3(h_W, i, j, k) + A3(h_E, i, j, k)) - 2.0 * A3(h, i, j, k);
                    h_avg = A3(h_E, i, j, k) + CFL * (0.5 * (A3(h_W, i, j, k) - A3(h_E, i, j, k)) + curv_3 * (CFL - 1.5));
                    h_marg = A3(h_E, i, j, k) + CFL * ((A3(h_W, i, j, k) - A3(h_E, i, j, k)) + 3.0 * curv_3 * (CFL - 1.0));
                } else if (uu < 0.0) {
                    if (m->vol_CFL) CFL = (-uu * dt) * (A2(m->dy_Cu, I, j) * A2(m->IareaT, i + 1, j));
                    else CFL = -(uu * dt * A2(m->IdxT, i + 1, j));
                    curv_3 = (A3(h_W, i + 1, j, k) + A3(h_E, i + 1, j, k)) - 2.0 * A3(h, i + 1, j, k);
                    h_avg = A3(h_W, i + 1, j, k) + CFL * (0.5 * (A3(h_E, i + 1, j, k) - A3(h_W, i + 1, j, k)) + curv_3 * (CFL - 1.5));
                    h_marg = A3(h_W, i + 1, j, k) + CFL * ((A3(h_E, i + 1, j, k) - A3(h_W, i + 1, j, k)) +
                                                          3.0 * curv_3 * (CFL - 1.0));
                } else {
                    h_avg = 0.5 * (A3(h_W, i + 1, j, k) + A3(h_E, i, j, k));
                    h_marg = 0.5 * (A3(h_W, i + 1, j, k) + A3(h_E, i, j, k));
                }
                A3(h_u, I, j, k) = m->marginal_faces ? h_marg : h_avg;
            }
    /* visc_rem_u is present in the driver's call (:503-505) */
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int I = m->isc - 1; I <= m->iec; ++I)
                A3(h_u, I, j, k) = A3(h_u, I, j, k) * (A3(m->visc_rem_u, I, j, k) * A3(m->por_face_areaU, I, j, k));
}

/* zonal_mass_flux, mom6_continuity_serial.F90:218-513, as called by
 * continuity_PPM with every optional argument present. */
static void zonal_mass_flux(om_model *m, const double *u, const double *h_in, const double *h_W,
                            const double *h_E, double *uh, double dt)
{
    const int n1 = m->n1, nz = m->nk;
    const int ish = m->isc, ieh = m->iec, jsh = m->jsc, jeh = m->jec;
    double *duhdu = calloc((size_t)n1 * nz, sizeof(double));
    double *visc_rem = calloc((size_t)n1 * nz, sizeof(double));
    double *du = calloc(n1, sizeof(double)), *du_min_CFL = calloc(n1, sizeof(double));
    double *du_max_CFL = calloc(n1, sizeof(double)), *duhdu_tot_0 = calloc(n1, sizeof(double));
    double *uh_tot_0 = calloc(n1, sizeof(double)), *visc_rem_max = calloc(n1, sizeof(double));
    int *do_I = calloc(n1, sizeof(int));
    const double *visc_rem_u = m->visc_rem_u, *por = m->por_face_areaU;

    for (size_t q = 0; q < m->N2; ++q) m->du_cor[q] = 0.0;  /* du_cor(:,:) = 0 (:287) */
    double CFL_dt = m->CFL_limit_adjust / dt;
    const double I_dt = 1.0 / dt;
    if (m->aggress_adjust) CFL_dt = I_dt;

    for (int j = jsh; j <= jeh; ++j) {
        for (int I = ish - 1; I <= ieh; ++I) V1(do_I, I) = 1;
        for (int k = 1; k <= nz; ++k) {
            for (int I = ish - 1; I <= ieh; ++I) R2(visc_rem, I, k) = A3(visc_rem_u, I, j, k);
            zonal_flux_layer(m, ROW(u, j, k), ROW(h_in, j, k), ROW(h_W, j, k), ROW(h_E, j, k),
                             ROW(uh, j, k), &R2(duhdu, 1, k), &R2(visc_rem, 1, k), dt, j, ish, ieh,
                             do_I, ROW(por, j, k));
        }
        if (m->use_visc_rem_max) {
            for (int I = 1; I <= n1; ++I) V1(visc_rem_max, I) = 0.0;
            for (int k = 1; k <= nz; ++k)
                for (int I = ish - 1; I <= ieh; ++I)
                    V1(visc_rem_max, I) = dmax(V1(visc_rem_max, I), R2(visc_rem, I, k));
        } else {
            for (int I = 1; I <= n1; ++I) V1(visc_rem_max, I) = 1.0;
        }
        for (int I = ish - 1; I <= ieh; ++I) {
            const int i = I;
            double I_vrm = 0.0, dx_W, dx_E;
            if (V1(visc_rem_max, I) > 0.0) I_vrm = 1.0 / V1(visc_rem_max, I);
            if (m->vol_CFL) {
                dx_W = ratio_max(A2(m->areaT, i, j), A2(m->dy_Cu, I, j), 1000.0 * A2(m->dxT, i, j));
                dx_E = ratio_max(A2(m->areaT, i + 1, j), A2(m->dy_Cu, I, j), 1000.0 * A2(m->dxT, i + 1, j));
            } else {
                dx_W = A2(m->dxT, i, j);
                dx_E = A2(m->dxT, i + 1, j);
            }
            V1(du_max_CFL, I) = 2.0 * (CFL_dt * dx_W) * I_vrm;
            V1(du_min_CFL, I) = -(2.0 * (CFL_dt * dx_E) * I_vrm);
            V1(uh_tot_0, I) = 0.0;
            V1(duhdu_tot_0, I) = 0.0;
        }
        for (int k = 1; k <= nz; ++k)
            for (int I = ish - 1; I <= ieh; ++I) {
                V1(duhdu_tot_0, I) = V1(duhdu_tot_0, I) + R2(duhdu, I, k);
                V1(uh_tot_0, I) = V1(uh_tot_0, I) + A3(uh, I, j, k);
            }
        for (int k = 1; k <= nz; ++k)
            for (int I = ish - 1; I <= ieh; ++I) {
                const int i = I;
                double dx_W, dx_E;
                if (m->vol_CFL) {
                    dx_W = ratio_max(A2(m->areaT, i, j), A2(m->dy_Cu, I, j), 1000.0 * A2(m->dxT, i, j));
                    dx_E = ratio_max(A2(m->areaT, i + 1, j), A2(m->dy_Cu, I, j), 1000.0 * A2(m->dxT, i + 1, j));
                } else {
                    dx_W = A2(m->dxT, i, j);
                    dx_E = A2(m->dxT, i + 1, j);
                }
                const double uu = A3(u, I, j, k), vr = R2(visc_rem, I, k);
                if (m->aggress_adjust) {
                    double du_lim = 0.499 * ((dx_W * I_dt - uu) + dmin(0.0, A3(u, I - 1, j, k)));
                    if (V1(du_max_CFL, I) * vr > du_lim) V1(du_max_CFL, I) = du_lim / vr;
                    du_lim = 0.499 * ((-dx_E * I_dt - uu) + dmax(0.0, A3(u, I + 1, j, k)));
                    if (V1(du_min_CFL, I) * vr < du_lim) V1(du_min_CFL, I) = du_lim / vr;
                } else {
                    if (V1(du_max_CFL, I) * vr > dx_W * CFL_dt - uu * A2(m->mask2dCu, I, j))
                        V1(du_max_CFL, I) = (dx_W * CFL_dt - uu) / vr;
                    if (V1(du_min_CFL, I) * vr < -dx_E * CFL_dt - uu * A2(m->mask2dCu, I, j))
                        V1(du_min_CFL, I) = -(dx_E * CFL_dt + uu) / vr;
                }
            }
        for (int I = ish - 1; I <= ieh; ++I) {
            V1(du_max_CFL, I) = dmax(V1(du_max_CFL, I), 0.0);
            V1(du_min_CFL, I) = dmin(V1(du_min_CFL, I), 0.0);
        }
        for (int I = ish - 1; I <= ieh; ++I) V1(do_I, I) = 1;

        zonal_flux_adjust(m, u, h_in, h_W, h_E, ROW(m->uhbt, j, 1), uh_tot_0, duhdu_tot_0, du,
                          du_max_CFL, du_min_CFL, dt, visc_rem, j, ish, ieh, do_I, por, uh);
        for (int k = 1; k <= nz; ++k)
            for (int I = ish - 1; I <= ieh; ++I)
                A3(m->u_cor, I, j, k) = A3(u, I, j, k) + V1(du, I) * R2(visc_rem, I, k);
        for (int I = ish - 1; I <= ieh; ++I) A2(m->du_cor, I, j) = V1(du, I);

        set_zonal_BT_cont(m, u, h_in, h_W, h_E, uh_tot_0, duhdu_tot_0, du_max_CFL, du_min_CFL, dt,
                          visc_rem, visc_rem_max, j, ish, ieh, do_I, por);
    }
    zonal_flux_thickness(m, m->u_cor, h_in, h_W, h_E, m->BT_h_u, dt);

    free(duhdu); free(visc_rem); free(du); free(du_min_CFL); free(du_max_CFL);
    free(duhdu_tot_0); free(uh_tot_0); free(visc_rem_max); free(do_I);
}

/* continuity_zonal_convergence, mom6_continuity_serial.F90:148-176 (hin present) */
static void continuity_zonal_convergence(om_model *m, double *h, const double *uh, double dt,
                                         const double *hin)
{
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec; ++i)
                A3(h, i, j, k) = dmax(A3(hin, i, j, k) - dt * A2(m->IareaT, i, j) *
                                                            (A3(uh, i, j, k) - A3(uh, i - 1, j, k)), 0.0);
}

/* continuity_PPM, mom6_continuity_serial.F90:100-145 */
void om_continuity_ppm(om_model *m, const double *u, const double *hin, double *h, double *uh, double dt)
{
    zonal_edge_thickness(m, hin, m->cont_hW, m->cont_hE);
    zonal_mass_flux(m, u, hin, m->cont_hW, m->cont_hE, uh, dt);
    continuity_zonal_convergence(m, h, uh, dt, hin);
}

/* Meridional pass (no counterpart in the reference, whose continuity is
 * zonal only).  MOM6's meridional mass flux is the zonal algorithm with
 * (i, u, dx, dy_Cu, IdxT, dxT, dxCu, mask2dCu, h_W/h_E) exchanged for
 * (j, v, dy, dx_Cv, IdyT, dyT, dyCv, mask2dCv, h_S/h_N); here it is exactly
 * that: the zonal pass above, pinned to the reference, applied to the
 * transposed grid and state (a second model with ni and nj exchanged). */
static void transpose3(const om_model *m, const double *a, om_model *t, double *b, int nk)
{
    for (int k = 1; k <= nk; ++k)
        for (int j = m->jsd; j <= m->jed; ++j)
            for (int i = m->isd; i <= m->ied; ++i)
                b[(size_t)(k - 1) * t->N2 + (size_t)(i - t->jsd) * t->n1 + (size_t)(j - t->isd)] =
                    a[(size_t)(k - 1) * m->N2 + (size_t)(j - m->jsd) * m->n1 + (size_t)(i - m->isd)];
}

static void transpose3_back(const om_model *t, const double *b, om_model *m, double *a, int nk)
{
    for (int k = 1; k <= nk; ++k)
        for (int j = m->jsd; j <= m->jed; ++j)
            for (int i = m->isd; i <= m->ied; ++i)
                a[(size_t)(k - 1) * m->N2 + (size_t)(j - m->jsd) * m->n1 + (size_t)(i - m->isd)] =
                    b[(size_t)(k - 1) * t->N2 + (size_t)(i - t->jsd) * t->n1 + (size_t)(j - t->isd)];
}

static om_model *transposed_model(om_model *m)
{
    if (m->mt) return m->mt;
    om_config c = m->c;
    c.ni = m->c.nj;
    c.nj = m->c.ni;
    om_model *t = om_create(&c);
    /* the y metrics of m become the x metrics of t */
    transpose3(m, m->IareaT, t, t->IareaT, 1);
    transpose3(m, m->IdyT, t, t->IdxT, 1);
    transpose3(m, m->dyT, t, t->dxT, 1);
    transpose3(m, m->dxCv, t, t->dy_Cu, 1);
    transpose3(m, m->dyCv, t, t->dxCu, 1);
    transpose3(m, m->mask2dT, t, t->mask2dT, 1);
    transpose3(m, m->mask2dCv, t, t->mask2dCu, 1);
    t->monotonic = m->monotonic; t->upwind_1st = m->upwind_1st; t->simple_2nd = m->simple_2nd;
    m->mt = t;
    return t;
}

void om_continuity_ppm_y(om_model *m, const double *v, const double *hin, double *h, double *vh, double dt)
{
    om_model *t = transposed_model(m);
    const int nk = m->nk;
    double *vT = (double *)malloc(t->N3 * sizeof(double)), *hinT = (double *)malloc(t->N3 * sizeof(double));
    double *hT = (double *)malloc(t->N3 * sizeof(double)), *vhT = (double *)malloc(t->N3 * sizeof(double));
    transpose3(m, v, t, vT, nk);
    transpose3(m, hin, t, hinT, nk);
    transpose3(m, h, t, hT, nk);
    transpose3(m, vh, t, vhT, nk);
    transpose3(m, m->vhbt, t, t->uhbt, 1);
    om_continuity_ppm(t, vT, hinT, hT, vhT, dt);
    transpose3_back(t, hT, m, h, nk);
    transpose3_back(t, vhT, m, vh, nk);
    transpose3_back(t, t->u_cor, m, m->v_cor, nk);
    transpose3_back(t, t->du_cor, m, m->dv_cor, 1);
    transpose3_back(t, t->FA_u_EE, m, m->FA_v_NN, 1);
    transpose3_back(t, t->FA_u_E0, m, m->FA_v_N0, 1);
    transpose3_back(t, t->FA_u_W0, m, m->FA_v_S0, 1);
    transpose3_back(t, t->FA_u_WW, m, m->FA_v_SS, 1);
    transpose3_back(t, t->uBT_WW, m, m->vBT_SS, 1);
    transpose3_back(t, t->uBT_EE, m, m->vBT_NN, 1);
    transpose3_back(t, t->BT_h_u, m, m->BT_h_v, nk);
    free(vT); free(hinT); free(hT); free(vhT);
}

/* ------------------------------------------------------------ coriolis --- */
/* coriolis_init Area_q, mom6_coriolis_serial.F90:74-80 */
static void coriolis_init(om_model *m)
{
    for (int j = m->jsd; j <= m->jed - 1; ++j)
        for (int i = m->isd; i <= m->ied - 1; ++i)
            A2(m->cor_Area_q, i, j) = (A2(m->areaT, i, j) + A2(m->areaT, i + 1, j + 1)) +
                                      (A2(m->areaT, i + 1, j) + A2(m->areaT, i, j + 1));
}

/* CorAdCalc, mom6_coriolis_serial.F90:112-252 */
void om_coradcalc(om_model *m, const double *u, const double *v, const double *h, const double *uh,
                  const double *vh, double *CAu, double *CAv)
{
    const int is = m->isc, ie = m->iec, js = m->jsc, je = m->jec, nz = m->nk;
    const double vol_neglect = 1.0e-20;
    const double C1_12 = 1.0 / 12.0, C1_24 = 1.0 / 24.0;
    const int scheme = m->c.coriolis_scheme;
    for (int k = 1; k <= nz; ++k) {
        for (int j = js - 1; j <= je; ++j)
            for (int i = is - 1; i <= ie; ++i) {
                A2(m->cor_dvdx, i, j) = (A3(v, i + 1, j, k) * A2(m->dyCv, i + 1, j)) - (A3(v, i, j, k) * A2(m->dyCv, i, j));
                A2(m->cor_dudy, i, j) = (A3(u, i, j + 1, k) * A2(m->dxCu, i, j + 1)) - (A3(u, i, j, k) * A2(m->dxCu, i, j));
            }
        for (int j = js - 1; j <= je; ++j)
            for (int i = is; i <= ie + 1; ++i)
                A2(m->cor_hArea_v, i, j) = 0.5 * ((A2(m->areaT, i, j) * A3(h, i, j, k)) + (A2(m->areaT, i, j + 1) * A3(h, i, j + 1, k)));
        for (int j = js; j <= je + 1; ++j)
            for (int i = is - 1; i <= ie; ++i)
                A2(m->cor_hArea_u, i, j) = 0.5 * ((A2(m->areaT, i, j) * A3(h, i, j, k)) + (A2(m->areaT, i + 1, j) * A3(h, i + 1, j, k)));
        for (int j = js - 1; j <= je; ++j)
            for (int i = is - 1; i <= ie; ++i) {
                A2(m->cor_rel_vort, i, j) = A2(m->mask2dBu, i, j) * (A2(m->cor_dvdx, i, j) - A2(m->cor_dudy, i, j)) * A2(m->IareaBu, i, j);
                A2(m->cor_abs_vort, i, j) = A2(m->CoriolisBu, i, j) + A2(m->cor_rel_vort, i, j);
            }
        for (int j = js - 1; j <= je; ++j)
            for (int i = is - 1; i <= ie; ++i) {
                const double hArea_q = (A2(m->cor_hArea_u, i, j) + A2(m->cor_hArea_u, i, j + 1)) +
                                       (A2(m->cor_hArea_v, i, j) + A2(m->cor_hArea_v, i + 1, j));
                A2(m->cor_Ih_q, i, j) = A2(m->cor_Area_q, i, j) / (hArea_q + vol_neglect);
                A2(m->cor_q, i, j) = A2(m->cor_abs_vort, i, j) * A2(m->cor_Ih_q, i, j);
            }
        double *q = m->cor_q, *a = m->cor_a, *b = m->cor_b, *c = m->cor_c, *d = m->cor_d;
        if (scheme == 2) {
            for (int j = js; j <= je; ++j)
                for (int i = is - 1; i <= ie; ++i) {
                    A2(a, i, j) = (A2(q, i, j) + (A2(q, i + 1, j) + A2(q, i, j - 1))) * C1_12;
                    A2(d, i, j) = ((A2(q, i, j) + A2(q, i + 1, j - 1)) + A2(q, i, j - 1)) * C1_12;
                }
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    A2(b, i, j) = (A2(q, i, j) + (A2(q, i - 1, j) + A2(q, i, j - 1))) * C1_12;
                    A2(c, i, j) = ((A2(q, i, j) + A2(q, i - 1, j - 1)) + A2(q, i, j - 1)) * C1_12;
                }
        } else if (scheme == 3) {
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    A2(a, i - 1, j) = (2.0 * (A2(q, i, j) + A2(q, i - 1, j - 1)) + (A2(q, i - 1, j) + A2(q, i, j - 1))) * C1_24;
                    A2(d, i - 1, j) = ((A2(q, i, j) + A2(q, i - 1, j - 1)) + 2.0 * (A2(q, i - 1, j) + A2(q, i, j - 1))) * C1_24;
                    A2(b, i, j) = ((A2(q, i, j) + A2(q, i - 1, j - 1)) + 2.0 * (A2(q, i - 1, j) + A2(q, i, j - 1))) * C1_24;
                    A2(c, i, j) = (2.0 * (A2(q, i, j) + A2(q, i - 1, j - 1)) + (A2(q, i - 1, j) + A2(q, i, j - 1))) * C1_24;
                }
        }
        for (int j = js; j <= je; ++j)
            for (int i = is; i <= ie; ++i) {
                const double uc = A3(u, i, j, k), uw = A3(u, i - 1, j, k);
                const double vc = A3(v, i, j, k), vs = A3(v, i, j - 1, k);
                A2(m->cor_KE, i, j) = 0.25 * ((A2(m->dyCu, i, j) * (uc * uc) + A2(m->dyCu, i - 1, j) * (uw * uw)) +
                                              (A2(m->dxCv, i, j) * (vc * vc) + A2(m->dxCv, i, j - 1) * (vs * vs))) /
                                      A2(m->areaT, i, j);
            }
        double *KE = m->cor_KE;
        if (scheme == 1) {
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie - 1; ++i) {
                    const double KEx = (A2(KE, i + 1, j) - A2(KE, i, j)) * A2(m->IdxCu, i, j);
                    A3(CAu, i, j, k) = 0.25 * ((A2(q, i, j) * (A3(vh, i + 1, j, k) + A3(vh, i, j, k))) +
                                               (A2(q, i, j - 1) * (A3(vh, i, j - 1, k) + A3(vh, i + 1, j - 1, k)))) *
                                           A2(m->IdxCu, i, j) - KEx;
                }
            for (int j = js; j <= je - 1; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double KEy = (A2(KE, i, j + 1) - A2(KE, i, j)) * A2(m->IdyCv, i, j);
                    A3(CAv, i, j, k) = -(0.25 * ((A2(q, i - 1, j) * (A3(uh, i - 1, j, k) + A3(uh, i - 1, j + 1, k))) +
                                                 (A2(q, i, j) * (A3(uh, i, j, k) + A3(uh, i, j + 1, k)))) *
                                         A2(m->IdyCv, i, j)) - KEy;
                }
        } else {
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie - 1; ++i) {
                    const double KEx = (A2(KE, i + 1, j) - A2(KE, i, j)) * A2(m->IdxCu, i, j);
                    A3(CAu, i, j, k) = (((A2(a, i, j) * A3(vh, i + 1, j, k)) + (A2(c, i, j) * A3(vh, i, j - 1, k))) +
                                        ((A2(b, i, j) * A3(vh, i, j, k)) + (A2(d, i, j) * A3(vh, i + 1, j - 1, k)))) *
                                           A2(m->IdxCu, i, j) - KEx;
                }
            for (int j = js; j <= je - 1; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double KEy = (A2(KE, i, j + 1) - A2(KE, i, j)) * A2(m->IdyCv, i, j);
                    A3(CAv, i, j, k) = -((((A2(a, i - 1, j) * A3(uh, i - 1, j, k)) + (A2(c, i, j + 1) * A3(uh, i, j + 1, k))) +
                                          ((A2(b, i, j) * A3(uh, i, j, k)) + (A2(d, i - 1, j + 1) * A3(uh, i - 1, j + 1, k)))) *
                                         A2(m->IdyCv, i, j)) - KEy;
                }
        }
    }
}

/* ----------------------------------------------------------- barotropic --- */
/* barotropic_init, mom6_barotropic_serial.F90:63-137 (plus wall masking of
 * Datv: closed v faces carry no barotropic transport) */
static void barotropic_init(om_model *m)
{
    const double depth = 4000.0;
    m->bt_nstep = m->c.bt_nsteps;
    m->bt_dtbt = m->c.dt / (double)m->bt_nstep;
    m->bt_bebt = 0.2;
    m->bt_dgeo_de = 1.0;
    for (int j = m->jsd; j <= m->jed; ++j)
        for (int i = m->isd; i <= m->ied; ++i) {
            A2(m->bt_Datu, i, j) = depth * A2(m->dyCu, i, j);
            A2(m->bt_Datv, i, j) = depth * A2(m->dxCv, i, j);
            if (m->c.bc_y == OM_BC_WALL) A2(m->bt_Datv, i, j) = A2(m->bt_Datv, i, j) * A2(m->mask2dCv, i, j);
            A2(m->bt_gtot_E, i, j) = G_EARTH;
            A2(m->bt_gtot_W, i, j) = G_EARTH;
            A2(m->bt_gtot_N, i, j) = G_EARTH;
            A2(m->bt_gtot_S, i, j) = G_EARTH;
            A2(m->bt_rem_u, i, j) = 0.999;
            A2(m->bt_rem_v, i, j) = 0.999;
        }
    for (int j = m->jsd; j <= m->jed; ++j)
        for (int i = m->isd; i <= m->ied; ++i) {
            const double f0 = A2(m->CoriolisBu, i, j);
            const size_t o = 4 * ((size_t)(j - 1) * m->n1 + (size_t)(i - 1));
            for (int c = 0; c < 4; ++c) {
                m->bt_f_4_u[o + c] = 0.25 * f0;
                m->bt_f_4_v[o + c] = 0.25 * f0;
            }
        }
}
#define F4(a, c, i, j) ((a)[4 * ((size_t)((j) - 1) * m->n1 + (size_t)((i) - 1)) + (c) - 1])

/* update_u / update_v, mom6_barotropic_serial.F90:307-356 */
static void bt_update_u(om_model *m, int is, int ie, int js, int je)
{
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_Cor_u, i, j) = (((F4(m->bt_f_4_u, 4, i, j) * A2(m->bt_vbt, i + 1, j)) + (F4(m->bt_f_4_u, 1, i, j) * A2(m->bt_vbt, i, j - 1))) +
                                     ((F4(m->bt_f_4_u, 3, i, j) * A2(m->bt_vbt, i, j)) + (F4(m->bt_f_4_u, 2, i, j) * A2(m->bt_vbt, i + 1, j - 1))));
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_ubt, i, j) = A2(m->bt_rem_u, i, j) * (A2(m->bt_ubt, i, j) + m->bt_dtbt * (A2(m->bt_Cor_u, i, j) + A2(m->bt_PFu, i, j)));
}
static void bt_update_v(om_model *m, int is, int ie, int js, int je)
{
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_Cor_v, i, j) = -1.0 * (((F4(m->bt_f_4_v, 1, i, j) * A2(m->bt_ubt, i - 1, j)) + (F4(m->bt_f_4_v, 4, i, j) * A2(m->bt_ubt, i, j + 1))) +
                                            ((F4(m->bt_f_4_v, 2, i, j) * A2(m->bt_ubt, i, j)) + (F4(m->bt_f_4_v, 3, i, j) * A2(m->bt_ubt, i - 1, j + 1))));
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_vbt, i, j) = A2(m->bt_rem_v, i, j) * (A2(m->bt_vbt, i, j) + m->bt_dtbt * (A2(m->bt_Cor_v, i, j) + A2(m->bt_PFv, i, j)));
}

/* One substep: btstep loop body (:210-293) or btstep_do_step (:379-474) when
 * extended = 1 (uhbt over I=is-1..ie, vhbt over J=js-1..je). */
static void bt_substep(om_model *m, int n, int extended)
{
    const int is = m->isc, ie = m->iec, js = m->jsc, je = m->jec;
    const double trans_wt1 = 1.0 + m->bt_bebt, trans_wt2 = -m->bt_bebt;
    const double inv_nstep = 1.0 / (double)m->bt_nstep, dtbt = m->bt_dtbt;
    for (int j = js; j <= je; ++j)
        for (int i = is - 1; i <= ie + 1; ++i) A2(m->bt_ubt_prev, i, j) = A2(m->bt_ubt, i, j);
    for (int j = js - 1; j <= je + 1; ++j)
        for (int i = is; i <= ie; ++i) A2(m->bt_vbt_prev, i, j) = A2(m->bt_vbt, i, j);
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_eta_pred, i, j) = A2(m->bt_eta, i, j) + (dtbt * A2(m->IareaT, i, j)) *
                (((A2(m->bt_Datu, i - 1, j) * A2(m->bt_ubt, i - 1, j)) - (A2(m->bt_Datu, i, j) * A2(m->bt_ubt, i, j))) +
                 ((A2(m->bt_Datv, i, j - 1) * A2(m->bt_vbt, i, j - 1)) - (A2(m->bt_Datv, i, j) * A2(m->bt_vbt, i, j))));
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie - 1; ++i)
            A2(m->bt_PFu, i, j) = ((A2(m->bt_eta_pred, i, j) * A2(m->bt_gtot_E, i, j)) -
                                   (A2(m->bt_eta_pred, i + 1, j) * A2(m->bt_gtot_W, i + 1, j))) * m->bt_dgeo_de * A2(m->IdxCu, i, j);
    for (int j = js; j <= je - 1; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_PFv, i, j) = ((A2(m->bt_eta_pred, i, j) * A2(m->bt_gtot_N, i, j)) -
                                   (A2(m->bt_eta_pred, i, j + 1) * A2(m->bt_gtot_S, i, j + 1))) * m->bt_dgeo_de * A2(m->IdyCv, i, j);
    const int v_first = ((n + m->first_direction) % 2) == 1;
    if (v_first) {
        bt_update_v(m, is, ie, js, je - 1);
        bt_update_u(m, is, ie - 1, js, je);
    } else {
        bt_update_u(m, is, ie - 1, js, je);
        bt_update_v(m, is, ie, js, je - 1);
    }
    const int iu0 = extended ? is - 1 : is, iu1 = extended ? ie : ie - 1;
    const int jv0 = extended ? js - 1 : js, jv1 = extended ? je : je - 1;
    for (int j = js; j <= je; ++j)
        for (int i = iu0; i <= iu1; ++i)
            A2(m->bt_uhbt, i, j) = A2(m->bt_Datu, i, j) * (trans_wt1 * A2(m->bt_ubt, i, j) + trans_wt2 * A2(m->bt_ubt_prev, i, j));
    for (int j = jv0; j <= jv1; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_vhbt, i, j) = A2(m->bt_Datv, i, j) * (trans_wt1 * A2(m->bt_vbt, i, j) + trans_wt2 * A2(m->bt_vbt_prev, i, j));
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_eta, i, j) = A2(m->bt_eta, i, j) - dtbt * A2(m->IareaT, i, j) *
                ((A2(m->bt_uhbt, i, j) - A2(m->bt_uhbt, i - 1, j)) + (A2(m->bt_vhbt, i, j) - A2(m->bt_vhbt, i, j - 1)));
    for (int j = js; j <= je; ++j)
        for (int i = is; i <= ie - 1; ++i)
            A2(m->bt_ubt_av, i, j) = A2(m->bt_ubt_av, i, j) + A2(m->bt_ubt, i, j) * inv_nstep;
    for (int j = js; j <= je - 1; ++j)
        for (int i = is; i <= ie; ++i)
            A2(m->bt_vbt_av, i, j) = A2(m->bt_vbt_av, i, j) + A2(m->bt_vbt, i, j) * inv_nstep;
}

/* btstep (:178-304) for static halos; for periodic grids the split API
 * btstep_init_state / btstep_do_step / btstep_get_output (:359-493) with a
 * halo refill of ubt, vbt, eta after every substep but the last
 * (rk2_mpi_serial_driver.F90:297-306). */
void om_btstep(om_model *m, const double *eta_in, const double *ubt_in, const double *vbt_in,
               double *u_av, double *v_av, double *eta_av)
{
    for (size_t q = 0; q < m->N2; ++q) {
        m->bt_eta[q] = eta_in[q];
        m->bt_ubt[q] = ubt_in[q];
        m->bt_vbt[q] = vbt_in[q];
        m->bt_ubt_av[q] = 0.0;
        m->bt_vbt_av[q] = 0.0;
    }
    const int split = periodic(m);
    for (int n = 1; n <= m->bt_nstep; ++n) {
        bt_substep(m, n, split);
        if (split && n < m->bt_nstep) {
            halo_fill(m, m->bt_ubt, 1);
            halo_fill(m, m->bt_vbt, 1);
            halo_fill(m, m->bt_eta, 1);
        }
    }
    for (size_t q = 0; q < m->N2; ++q) {
        u_av[q] = m->bt_ubt_av[q];
        v_av[q] = m->bt_vbt_av[q];
        eta_av[q] = m->bt_eta[q];
    }
}

/* ------------------------------------------------------------ vert_visc --- */
static inline double botfn_of(double z2)
{
    return 1.0 / (1.0 + 0.09 * z2 * z2 * z2 * z2 * z2 * z2);
}

/* vert_visc_coef for one velocity column, mom6_vert_visc_serial.F90:155-223.
 * hA/hB are the two thickness columns (h(i) and h(i+1), or h(j) and h(j+1)). */
static void vv_coef_column(om_model *m, double *a, double *hvel_out, const double *vel,
                           const double *hA, const double *hB, size_t stride, double mask,
                           double *hvel, double *z_i)
{
    const int nz = m->nk;
    const double h_neglect = 1.0e-10;
    const double I_Hbbl = 1.0 / (m->vv_Hbbl + h_neglect);
    if (mask > 0.0) {
        z_i[nz] = 0.0;  /* z_i(nz+1) */
        for (int k = nz; k >= 1; --k) {
            const size_t o = (size_t)(k - 1) * stride;
            const double h_harm = 2.0 * hA[o] * hB[o] / (hA[o] + hB[o] + h_neglect);
            const double h_arith = 0.5 * (hB[o] + hA[o]);
            const double h_delta = hB[o] - hA[o];
            hvel[k - 1] = h_harm;
            if (vel[o] * h_delta < 0.0) {
                const double botfn = botfn_of(z_i[k]);
                hvel[k - 1] = (1.0 - botfn) * h_harm + botfn * h_arith;
            }
            z_i[k - 1] = z_i[k] + h_harm * I_Hbbl;
        }
        a[0] = 0.0;
        double z_top = 0.0;
        for (int K2 = 2; K2 <= nz; ++K2) {
            z_top = z_top + hvel[K2 - 2];
            double Kv_tot = m->vv_Kv;
            if (z_top < m->vv_Hmix) {
                const double topfn = 1.0 - z_top / m->vv_Hmix;
                Kv_tot = Kv_tot + (m->vv_Kv_ml - m->vv_Kv) * topfn;
            }
            const double botfn = botfn_of(z_i[K2 - 1]);
            Kv_tot = Kv_tot + m->vv_Kv_extra_bbl * botfn;
            const double h_shear = 0.5 * (hvel[K2 - 1] + hvel[K2 - 2] + h_neglect);
            a[(size_t)(K2 - 1) * stride] = Kv_tot / h_shear;
        }
        a[(size_t)nz * stride] = (m->vv_Kv + m->vv_Kv_extra_bbl) / (0.5 * hvel[nz - 1] + h_neglect);
        for (int k = 1; k <= nz; ++k) hvel_out[(size_t)(k - 1) * stride] = hvel[k - 1] + h_neglect;
    } else {
        for (int k = 1; k <= nz; ++k) hvel_out[(size_t)(k - 1) * stride] = h_neglect;
        for (int K2 = 1; K2 <= nz + 1; ++K2) a[(size_t)(K2 - 1) * stride] = 0.0;
    }
}

/* vert_visc_apply for one column, mom6_vert_visc_serial.F90:417-460 */
static double vv_apply_column(om_model *m, double *vel, const double *hvel, const double *a,
                              const double *Ray, size_t stride, double dt, double sfc_stress,
                              double *c1)
{
    const int nz = m->nk;
    double Rayv = Ray[0];
    double b_denom_1 = hvel[0] + dt * (Rayv + a[0]);
    double b1 = 1.0 / (b_denom_1 + dt * a[stride]);
    double d1 = b_denom_1 * b1;
    vel[0] = b1 * (hvel[0] * vel[0] + sfc_stress);
    for (int k = 2; k <= nz; ++k) {
        const size_t o = (size_t)(k - 1) * stride;
        Rayv = Ray[o];
        c1[k - 1] = dt * a[o] * b1;
        b_denom_1 = hvel[o] + dt * (Rayv + a[o] * d1);
        b1 = 1.0 / (b_denom_1 + dt * a[o + stride]);
        d1 = b_denom_1 * b1;
        vel[o] = (hvel[o] * vel[o] + dt * a[o] * vel[o - stride]) * b1;
    }
    for (int k = nz - 1; k >= 1; --k) {
        const size_t o = (size_t)(k - 1) * stride;
        vel[o] = vel[o] + c1[k] * vel[o + stride];
    }
    double tbot = RHO_0 * vel[(size_t)(nz - 1) * stride] * a[(size_t)nz * stride];
    for (int k = 1; k <= nz; ++k) {
        const size_t o = (size_t)(k - 1) * stride;
        tbot = tbot + RHO_0 * Ray[o] * vel[o];
    }
    return tbot;
}

/* vert_visc_coef + vert_visc_apply (rk2_serial_driver.F90:276-277, 356-357);
 * the coupling arrays a_u/a_v have nz+1 levels. */
void om_vert_visc(om_model *m, double *u, double *v, const double *h, double dt)
{
    const int is = m->isc, ie = m->iec, js = m->jsc, je = m->jec, nz = m->nk;
    const size_t S = m->N2;   /* k stride of a 3D (or nz+1) array */
    double *hvel = calloc(nz + 1, sizeof(double)), *z_i = calloc(nz + 2, sizeof(double));
    double *c1 = calloc(nz + 1, sizeof(double));
    /* coefficients */
    for (int j = js; j <= je; ++j)
        for (int i = is - 1; i <= ie; ++i)
            vv_coef_column(m, &A3(m->vv_a_u, i, j, 1), &A3(m->vv_h_u, i, j, 1), &A3(u, i, j, 1),
                           &A3(h, i, j, 1), &A3(h, i + 1, j, 1), S, A2(m->mask2dCu, i, j), hvel, z_i);
    for (int j = js - 1; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            vv_coef_column(m, &A3(m->vv_a_v, i, j, 1), &A3(m->vv_h_v, i, j, 1), &A3(v, i, j, 1),
                           &A3(h, i, j, 1), &A3(h, i, j + 1, 1), S, A2(m->mask2dCv, i, j), hvel, z_i);
    /* apply */
    const double dt_Rho0 = dt / RHO_0;
    for (int j = js; j <= je; ++j)
        for (int i = is - 1; i <= ie; ++i)
            if (A2(m->mask2dCu, i, j) > 0.0) {
                const double sfc = dt_Rho0 * A2(m->taux, i, j) * A2(m->mask2dCu, i, j);
                A2(m->vv_taux_bot, i, j) = vv_apply_column(m, &A3(u, i, j, 1), &A3(m->vv_h_u, i, j, 1),
                                                           &A3(m->vv_a_u, i, j, 1), &A3(m->Ray_u, i, j, 1),
                                                           S, dt, sfc, c1);
            }
    for (int j = js - 1; j <= je; ++j)
        for (int i = is; i <= ie; ++i)
            if (A2(m->mask2dCv, i, j) > 0.0) {
                const double sfc = dt_Rho0 * A2(m->tauy, i, j) * A2(m->mask2dCv, i, j);
                A2(m->vv_tauy_bot, i, j) = vv_apply_column(m, &A3(v, i, j, 1), &A3(m->vv_h_v, i, j, 1),
                                                           &A3(m->vv_a_v, i, j, 1), &A3(m->Ray_v, i, j, 1),
                                                           S, dt, sfc, c1);
            }
    free(hvel); free(z_i); free(c1);
}

/* ------------------------------------------------------------ hor_visc --- */
/* hor_visc_init, mom6_hor_visc_serial.F90:160-286 */
static void hor_visc_init(om_model *m)
{
    const om_config *c = &m->c;
    m->hv_Laplacian = c->laplacian;
    m->hv_biharmonic = c->biharmonic;
    m->hv_Smag_Kh = c->smagorinsky && c->laplacian;
    m->hv_Smag_Ah = c->smagorinsky && c->biharmonic;
    m->hv_no_slip = c->no_slip;
    m->hv_Leith_Kh = c->leith;                        /* :184-187 */
    m->hv_Leith_Ah = c->leith && c->biharmonic;
    m->hv_bound_Kh = c->bound_Kh || c->better_bound;
    m->hv_bound_Ah = c->better_bound;
    m->hv_better_bound_Kh = c->better_bound;
    m->hv_better_bound_Ah = c->better_bound;
    m->hv_use_land_mask = 1;
    m->hv_Kh_bg_min = 0.0;
    m->hv_h_neglect = dmax(1.0e-10, 1.0e-3);
    const double Smag_Lap_const = 0.15, Smag_bi_const = 0.06, Leith_const = 1.0;   /* :55-57 */
    for (int j = m->jsd; j <= m->jed; ++j)
        for (int i = m->isd; i <= m->ied; ++i) {
            A2(m->hv_dx2h, i, j) = A2(m->dxT, i, j) * A2(m->dxT, i, j);
            A2(m->hv_dy2h, i, j) = A2(m->dyT, i, j) * A2(m->dyT, i, j);
            A2(m->hv_dx2q, i, j) = A2(m->dxBu, i, j) * A2(m->dxBu, i, j);
            A2(m->hv_dy2q, i, j) = A2(m->dyBu, i, j) * A2(m->dyBu, i, j);
            A2(m->hv_DX_dyT, i, j) = A2(m->dxT, i, j) * A2(m->IdyT, i, j);
            A2(m->hv_DY_dxT, i, j) = A2(m->dyT, i, j) * A2(m->IdxT, i, j);
            A2(m->hv_DX_dyBu, i, j) = A2(m->dxBu, i, j) * A2(m->IdyBu, i, j);
            A2(m->hv_DY_dxBu, i, j) = A2(m->dyBu, i, j) * A2(m->IdxBu, i, j);
            A2(m->hv_Idx2dyCu, i, j) = A2(m->IdxCu, i, j) * A2(m->IdxCu, i, j) * A2(m->IdyCu, i, j);
            A2(m->hv_Idxdy2u, i, j) = A2(m->IdxCu, i, j) * A2(m->IdyCu, i, j) * A2(m->IdyCu, i, j);
            A2(m->hv_Idx2dyCv, i, j) = A2(m->IdxCv, i, j) * A2(m->IdxCv, i, j) * A2(m->IdyCv, i, j);
            A2(m->hv_Idxdy2v, i, j) = A2(m->IdxCv, i, j) * A2(m->IdyCv, i, j) * A2(m->IdyCv, i, j);
            A2(m->hv_reduction_xx, i, j) = 1.0;
            A2(m->hv_reduction_xy, i, j) = 1.0;
            const double grid_sp_h2 = A2(m->areaT, i, j);
            const double grid_sp_h3 = sqrt(grid_sp_h2) * grid_sp_h2;
            const double grid_sp_h4 = grid_sp_h2 * grid_sp_h2;
            A2(m->hv_Laplac3_const_xx, i, j) = Leith_const * grid_sp_h3;
            A2(m->hv_Laplac3_const_xy, i, j) = Leith_const * grid_sp_h3;
            A2(m->hv_Biharm6_const_xx, i, j) = Leith_const * grid_sp_h3 * grid_sp_h3;
            A2(m->hv_Biharm6_const_xy, i, j) = Leith_const * grid_sp_h3 * grid_sp_h3;
            A2(m->hv_Kh_bg_xx, i, j) = c->Kh;
            A2(m->hv_Kh_bg_xy, i, j) = c->Kh;
            A2(m->hv_Ah_bg_xx, i, j) = c->Ah;
            A2(m->hv_Ah_bg_xy, i, j) = c->Ah;
            A2(m->hv_Kh_Max_xx, i, j) = 0.25 * grid_sp_h2 / 1.0;
            A2(m->hv_Kh_Max_xy, i, j) = 0.25 * grid_sp_h2 / 1.0;
            A2(m->hv_Ah_Max_xx, i, j) = 0.0625 * grid_sp_h4 / 1.0;
            A2(m->hv_Ah_Max_xy, i, j) = 0.0625 * grid_sp_h4 / 1.0;
            A2(m->hv_Laplac2_const_xx, i, j) = Smag_Lap_const * Smag_Lap_const * grid_sp_h2;
            A2(m->hv_Laplac2_const_xy, i, j) = Smag_Lap_const * Smag_Lap_const * grid_sp_h2;
            A2(m->hv_Biharm_const_xx, i, j) = Smag_bi_const * Smag_bi_const * grid_sp_h4;
            A2(m->hv_Biharm_const_xy, i, j) = Smag_bi_const * Smag_bi_const * grid_sp_h4;
        }
    for (size_t q = 0; q < m->N2; ++q) {   /* work arrays: zero, hrat/visc_bound_rem: one (:270-282) */
        m->hv_hrat_min[q] = 1.0;
        m->hv_visc_bound_rem[q] = 1.0;
    }
}

static inline double sq(double x) { return x * x; }

/* inv_PI3 = 1/(3.14159265358979_dp**3), inv_PI6 = 1/(3.14159265358979_dp**6)
 * (mom6_hor_visc_serial.F90:154-155); the integer powers by squaring, as the
 * compiled reference evaluates them (x**6 = (x*x * x*x) * x*x; a left-to-right
 * product differs in the last bit, which the Leith golden vectors show) */
#define PI_LEITH 3.14159265358979
#define INV_PI3 (1.0 / (PI_LEITH * PI_LEITH * PI_LEITH))
#define INV_PI6 (1.0 / (((PI_LEITH * PI_LEITH) * (PI_LEITH * PI_LEITH)) * (PI_LEITH * PI_LEITH)))

/* hor_visc, mom6_hor_visc_serial.F90:464-1109 (Laplacian, biharmonic,
 * Smagorinsky, Leith and the Kh/Ah bounds; FrictWork and modified Leith are not on the path) */
void om_hor_visc(om_model *m, const double *u, const double *v, const double *h, double *diffu, double *diffv)
{
    const int is = m->isc, ie = m->iec, js = m->jsc, je = m->jec, nz = m->nk;
    const int Isq = is - 1, Ieq = ie, Jsq = js - 1, Jeq = je;
    const double h_neglect = m->hv_h_neglect;
    for (size_t q = 0; q < m->N3; ++q) { diffu[q] = 0.0; diffv[q] = 0.0; }
    for (int k = 1; k <= nz; ++k) {
        for (int j = Jsq; j <= Jeq + 1; ++j)
            for (int i = Isq; i <= Ieq + 1; ++i) {
                A2(m->hv_dudx, i, j) = A2(m->hv_DY_dxT, i, j) * ((A2(m->IdyCu, i, j) * A3(u, i, j, k)) - (A2(m->IdyCu, i - 1, j) * A3(u, i - 1, j, k)));
                A2(m->hv_dvdy, i, j) = A2(m->hv_DX_dyT, i, j) * ((A2(m->IdxCv, i, j) * A3(v, i, j, k)) - (A2(m->IdxCv, i, j - 1) * A3(v, i, j - 1, k)));
            }
        for (int J = Jsq; J <= Jeq; ++J)
            for (int I = Isq; I <= Ieq; ++I) {
                A2(m->hv_dvdx, I, J) = A2(m->hv_DY_dxBu, I, J) * ((A3(v, I + 1, J, k) * A2(m->IdyCv, I + 1, J)) - (A3(v, I, J, k) * A2(m->IdyCv, I, J)));
                A2(m->hv_dudy, I, J) = A2(m->hv_DX_dyBu, I, J) * ((A3(u, I, J + 1, k) * A2(m->IdxCu, I, J + 1)) - (A3(u, I, J, k) * A2(m->IdxCu, I, J)));
            }
        for (int j = Jsq; j <= Jeq + 1; ++j)
            for (int i = Isq; i <= Ieq + 1; ++i) A2(m->hv_sh_xx, i, j) = A2(m->hv_dudx, i, j) - A2(m->hv_dvdy, i, j);
        for (int J = Jsq; J <= Jeq; ++J)
            for (int I = Isq; I <= Ieq; ++I)
                A2(m->hv_sh_xy, I, J) = (m->hv_no_slip ? (2.0 - A2(m->mask2dBu, I, J)) : A2(m->mask2dBu, I, J)) *
                                        (A2(m->hv_dvdx, I, J) + A2(m->hv_dudy, I, J));
        for (int j = js - 1; j <= je + 1; ++j)
            for (int I = Isq; I <= Ieq; ++I)
                A2(m->hv_h_u, I, j) = 0.5 * (A2(m->mask2dT, I, j) * A3(h, I, j, k) + A2(m->mask2dT, I + 1, j) * A3(h, I + 1, j, k));
        for (int J = Jsq; J <= Jeq; ++J)
            for (int i = is - 1; i <= ie + 1; ++i)
                A2(m->hv_h_v, i, J) = 0.5 * (A2(m->mask2dT, i, J) * A3(h, i, J, k) + A2(m->mask2dT, i, J + 1) * A3(h, i, J + 1, k));
        for (int J = Jsq; J <= Jeq; ++J)
            for (int I = Isq; I <= Ieq; ++I)
                A2(m->hv_hq, I, J) = 0.25 * ((A3(h, I, J, k) + A3(h, I + 1, J + 1, k)) + (A3(h, I + 1, J, k) + A3(h, I, J + 1, k)));
        if (m->hv_biharmonic) {
            for (int j = js; j <= je; ++j)
                for (int I = Isq; I <= Ieq; ++I)
                    A2(m->hv_Del2u, I, j) = A2(m->hv_Idx2dyCu, I, j) * ((A2(m->hv_dx2q, I, j) * A2(m->hv_sh_xy, I, j)) - (A2(m->hv_dx2q, I, j - 1) * A2(m->hv_sh_xy, I, j - 1))) +
                                            A2(m->hv_Idxdy2u, I, j) * ((A2(m->hv_dy2h, I + 1, j) * A2(m->hv_sh_xx, I + 1, j)) - (A2(m->hv_dy2h, I, j) * A2(m->hv_sh_xx, I, j)));
            for (int J = Jsq; J <= Jeq; ++J)
                for (int i = is; i <= ie; ++i)
                    A2(m->hv_Del2v, i, J) = A2(m->hv_Idxdy2v, i, J) * ((A2(m->hv_dy2q, i, J) * A2(m->hv_sh_xy, i, J)) - (A2(m->hv_dy2q, i - 1, J) * A2(m->hv_sh_xy, i - 1, J))) -
                                            A2(m->hv_Idx2dyCv, i, J) * ((A2(m->hv_dx2h, i, J + 1) * A2(m->hv_sh_xx, i, J + 1)) - (A2(m->hv_dx2h, i, J) * A2(m->hv_sh_xx, i, J)));
        }
        if (m->hv_Leith_Kh) {   /* Leith: vorticity and its gradients (:610-727; modified Leith off) */
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I)
                    A2(m->hv_vort_xy, I, J) = (m->hv_no_slip ? (2.0 - A2(m->mask2dBu, I, J)) : A2(m->mask2dBu, I, J)) *
                                              (A2(m->hv_dvdx, I, J) - A2(m->hv_dudy, I, J));
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double DY_dxBu_loc = A2(m->dyBu, i, j) * A2(m->IdxBu, i, j);
                    A2(m->hv_vort_xy_dx, i, j) = DY_dxBu_loc * ((A2(m->hv_vort_xy, i, j) * A2(m->IdyCu, i, j)) -
                                                                (A2(m->hv_vort_xy, i - 1, j) * A2(m->IdyCu, i - 1, j)));
                }
            for (int j = js; j <= je; ++j)
                for (int I = Isq; I <= Ieq; ++I) {
                    const double DX_dyBu_loc = A2(m->dxBu, I, j) * A2(m->IdyBu, I, j);
                    A2(m->hv_vort_xy_dy, I, j) = DX_dyBu_loc * ((A2(m->hv_vort_xy, I, j) * A2(m->IdxCv, I, j)) -
                                                                (A2(m->hv_vort_xy, I, j - 1) * A2(m->IdxCv, I, j - 1)));
                }
            if (m->hv_Leith_Ah)
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I)
                        A2(m->hv_Del2vort_q, I, J) =
                            A2(m->hv_DY_dxBu, I, J) * ((A2(m->hv_vort_xy_dx, I + 1, J) * A2(m->IdyCv, I + 1, J)) -
                                                       (A2(m->hv_vort_xy_dx, I, J) * A2(m->IdyCv, I, J))) +
                            A2(m->hv_DX_dyBu, I, J) * ((A2(m->hv_vort_xy_dy, I, J + 1) * A2(m->IdxCu, I, J + 1)) -
                                                       (A2(m->hv_vort_xy_dy, I, J) * A2(m->IdxCu, I, J)));
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double gx = 0.5 * (A2(m->hv_vort_xy_dx, i, j) + A2(m->hv_vort_xy_dx, i, j - 1));
                    const double gy = 0.5 * (A2(m->hv_vort_xy_dy, i, j) + A2(m->hv_vort_xy_dy, i - 1, j));
                    A2(m->hv_grad_vort_mag_h, i, j) = sqrt(sq(gx) + sq(gy));
                    A2(m->hv_vert_vort_mag, i, j) = A2(m->hv_grad_vort_mag_h, i, j) + 0.0;   /* + grad_div_mag_h = 0 */
                }
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) {
                    const double gx = 0.5 * (A2(m->hv_vort_xy_dx, I, J) + A2(m->hv_vort_xy_dx, I + 1, J));
                    const double gy = 0.5 * (A2(m->hv_vort_xy_dy, I, J) + A2(m->hv_vort_xy_dy, I, J + 1));
                    A2(m->hv_grad_vort_mag_q, I, J) = sqrt(sq(gx) + sq(gy));
                }
        }
        if (m->hv_Smag_Kh || m->hv_Smag_Ah)
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double sh_xx_sq = sq(A2(m->hv_sh_xx, i, j));
                    const double sh_xy_sq = 0.25 * ((sq(A2(m->hv_sh_xy, i - 1, j - 1)) + sq(A2(m->hv_sh_xy, i, j))) +
                                                    (sq(A2(m->hv_sh_xy, i - 1, j)) + sq(A2(m->hv_sh_xy, i, j - 1))));
                    A2(m->hv_Shear_mag, i, j) = sqrt(sh_xx_sq + sh_xy_sq);
                }
        if (m->hv_better_bound_Kh || m->hv_better_bound_Ah)
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    const double h_min = dmin(dmin(dmin(A2(m->hv_h_u, i, j), A2(m->hv_h_u, i - 1, j)), A2(m->hv_h_v, i, j)), A2(m->hv_h_v, i, j - 1));
                    A2(m->hv_hrat_min, i, j) = dmin(1.0, h_min / (A3(h, i, j, k) + h_neglect));
                }
        if (m->hv_Laplacian) {
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) A2(m->hv_Kh, i, j) = A2(m->hv_Kh_bg_xx, i, j);
            if (m->hv_Smag_Kh)
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i)
                        A2(m->hv_Kh, i, j) = dmax(A2(m->hv_Kh, i, j), A2(m->hv_Laplac2_const_xx, i, j) * A2(m->hv_Shear_mag, i, j));
            if (m->hv_Leith_Kh)
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i)
                        A2(m->hv_Kh, i, j) = dmax(A2(m->hv_Kh, i, j), A2(m->hv_Laplac3_const_xx, i, j) * A2(m->hv_vert_vort_mag, i, j) * INV_PI3);
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) A2(m->hv_Kh, i, j) = dmax(A2(m->hv_Kh, i, j), m->hv_Kh_bg_min);
            if (m->hv_better_bound_Kh && m->hv_better_bound_Ah) {
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i) {
                        A2(m->hv_visc_bound_rem, i, j) = 1.0;
                        const double Kh_max_here = A2(m->hv_hrat_min, i, j) * A2(m->hv_Kh_Max_xx, i, j);
                        if (A2(m->hv_Kh, i, j) >= Kh_max_here) {
                            A2(m->hv_visc_bound_rem, i, j) = 0.0;
                            A2(m->hv_Kh, i, j) = Kh_max_here;
                        } else if (A2(m->hv_Kh, i, j) > 0.0) {
                            A2(m->hv_visc_bound_rem, i, j) = 1.0 - A2(m->hv_Kh, i, j) / Kh_max_here;
                        }
                    }
            } else if (m->hv_better_bound_Kh) {
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i)
                        A2(m->hv_Kh, i, j) = dmin(A2(m->hv_Kh, i, j), A2(m->hv_hrat_min, i, j) * A2(m->hv_Kh_Max_xx, i, j));
            } else if (m->hv_bound_Kh) {
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i) A2(m->hv_Kh, i, j) = dmin(A2(m->hv_Kh, i, j), A2(m->hv_Kh_Max_xx, i, j));
            }
            for (int j = Jsq; j <= Jeq + 1; ++j)
                for (int i = Isq; i <= Ieq + 1; ++i) A2(m->hv_str_xx, i, j) = -(A2(m->hv_Kh, i, j) * A2(m->hv_sh_xx, i, j));
        } else {
            for (int j = Jsq; j <= Jeq + 1; ++j)
                for (int i = Isq; i <= Ieq + 1; ++i) A2(m->hv_str_xx, i, j) = 0.0;
        }
        if (m->hv_biharmonic) {
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) A2(m->hv_Ah, i, j) = A2(m->hv_Ah_bg_xx, i, j);
            if (m->hv_Smag_Ah)
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i)
                        A2(m->hv_Ah, i, j) = dmax(A2(m->hv_Ah, i, j), A2(m->hv_Biharm_const_xx, i, j) * A2(m->hv_Shear_mag, i, j));
            if (m->hv_Leith_Ah)
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i) {
                        const double Del2vort_h = 0.25 * ((A2(m->hv_Del2vort_q, i, j) + A2(m->hv_Del2vort_q, i - 1, j - 1)) +
                                                          (A2(m->hv_Del2vort_q, i - 1, j) + A2(m->hv_Del2vort_q, i, j - 1)));
                        A2(m->hv_Ah, i, j) = dmax(A2(m->hv_Ah, i, j), A2(m->hv_Biharm6_const_xx, i, j) * fabs(Del2vort_h) * INV_PI6);
                    }
            if (m->hv_better_bound_Ah) {
                if (m->hv_better_bound_Kh) {
                    for (int j = js; j <= je; ++j)
                        for (int i = is; i <= ie; ++i)
                            A2(m->hv_Ah, i, j) = dmin(A2(m->hv_Ah, i, j), A2(m->hv_visc_bound_rem, i, j) * A2(m->hv_hrat_min, i, j) * A2(m->hv_Ah_Max_xx, i, j));
                } else {
                    for (int j = js; j <= je; ++j)
                        for (int i = is; i <= ie; ++i)
                            A2(m->hv_Ah, i, j) = dmin(A2(m->hv_Ah, i, j), A2(m->hv_hrat_min, i, j) * A2(m->hv_Ah_Max_xx, i, j));
                }
            } else if (m->hv_bound_Ah) {
                for (int j = js; j <= je; ++j)
                    for (int i = is; i <= ie; ++i) A2(m->hv_Ah, i, j) = dmin(A2(m->hv_Ah, i, j), A2(m->hv_Ah_Max_xx, i, j));
            }
            for (int j = Jsq; j <= Jeq + 1; ++j)
                for (int i = Isq; i <= Ieq + 1; ++i) {
                    const double d_del2u = (A2(m->IdyCu, i, j) * A2(m->hv_Del2u, i, j)) - (A2(m->IdyCu, i - 1, j) * A2(m->hv_Del2u, i - 1, j));
                    const double d_del2v = (A2(m->IdxCv, i, j) * A2(m->hv_Del2v, i, j)) - (A2(m->IdxCv, i, j - 1) * A2(m->hv_Del2v, i, j - 1));
                    const double d_str = A2(m->hv_Ah, i, j) * ((A2(m->hv_DY_dxT, i, j) * d_del2u) - (A2(m->hv_DX_dyT, i, j) * d_del2v));
                    A2(m->hv_bhstr_xx, i, j) = d_str * (A3(h, i, j, k) * A2(m->hv_reduction_xx, i, j));
                    A2(m->hv_str_xx, i, j) = A2(m->hv_str_xx, i, j) + d_str;
                }
        }
        for (int j = Jsq; j <= Jeq + 1; ++j)
            for (int i = Isq; i <= Ieq + 1; ++i)
                A2(m->hv_str_xx, i, j) = A2(m->hv_str_xx, i, j) * (A3(h, i, j, k) * A2(m->hv_reduction_xx, i, j));
        if (m->hv_Laplacian) {
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) A2(m->hv_Kh, I, J) = A2(m->hv_Kh_bg_xy, I, J);
            if (m->hv_Smag_Kh)
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) {
                        const double sh_xx_sq = 0.25 * ((sq(A2(m->hv_sh_xx, I, J)) + sq(A2(m->hv_sh_xx, I + 1, J + 1))) +
                                                        (sq(A2(m->hv_sh_xx, I + 1, J)) + sq(A2(m->hv_sh_xx, I, J + 1))));
                        const double sh_xy_sq = sq(A2(m->hv_sh_xy, I, J));
                        A2(m->hv_Kh, I, J) = dmax(A2(m->hv_Kh, I, J), A2(m->hv_Laplac2_const_xy, I, J) * sqrt(sh_xx_sq + sh_xy_sq));
                    }
            if (m->hv_Leith_Kh)
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I)
                        A2(m->hv_Kh, I, J) = dmax(A2(m->hv_Kh, I, J), A2(m->hv_Laplac3_const_xy, I, J) * A2(m->hv_grad_vort_mag_q, I, J) * INV_PI3);
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) A2(m->hv_Kh, I, J) = dmax(A2(m->hv_Kh, I, J), m->hv_Kh_bg_min);
            if (m->hv_better_bound_Kh) {
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) {
                        const double h_min = 0.25 * ((A2(m->hv_hrat_min, I, J) + A2(m->hv_hrat_min, I + 1, J + 1)) +
                                                     (A2(m->hv_hrat_min, I + 1, J) + A2(m->hv_hrat_min, I, J + 1)));
                        A2(m->hv_Kh, I, J) = dmin(A2(m->hv_Kh, I, J), h_min * A2(m->hv_Kh_Max_xy, I, J));
                    }
            } else if (m->hv_bound_Kh) {
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) A2(m->hv_Kh, I, J) = dmin(A2(m->hv_Kh, I, J), A2(m->hv_Kh_Max_xy, I, J));
            }
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) A2(m->hv_str_xy, I, J) = -(A2(m->hv_Kh, I, J) * A2(m->hv_sh_xy, I, J));
        } else {
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) A2(m->hv_str_xy, I, J) = 0.0;
        }
        if (m->hv_biharmonic) {
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) A2(m->hv_Ah, I, J) = A2(m->hv_Ah_bg_xy, I, J);
            if (m->hv_Smag_Ah)
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) {
                        const double sh_xx_sq = 0.25 * ((sq(A2(m->hv_sh_xx, I, J)) + sq(A2(m->hv_sh_xx, I + 1, J + 1))) +
                                                        (sq(A2(m->hv_sh_xx, I + 1, J)) + sq(A2(m->hv_sh_xx, I, J + 1))));
                        const double sh_xy_sq = sq(A2(m->hv_sh_xy, I, J));
                        A2(m->hv_Ah, I, J) = dmax(A2(m->hv_Ah, I, J), A2(m->hv_Biharm_const_xy, I, J) * sqrt(sh_xx_sq + sh_xy_sq));
                    }
            if (m->hv_Leith_Ah)
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I)
                        A2(m->hv_Ah, I, J) = dmax(A2(m->hv_Ah, I, J), A2(m->hv_Biharm6_const_xy, I, J) * fabs(A2(m->hv_Del2vort_q, I, J)) * INV_PI6);
            if (m->hv_better_bound_Ah) {
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) {
                        const double h_min = 0.25 * ((A2(m->hv_hrat_min, I, J) + A2(m->hv_hrat_min, I + 1, J + 1)) +
                                                     (A2(m->hv_hrat_min, I + 1, J) + A2(m->hv_hrat_min, I, J + 1)));
                        A2(m->hv_Ah, I, J) = dmin(A2(m->hv_Ah, I, J), h_min * A2(m->hv_Ah_Max_xy, I, J));
                    }
            } else if (m->hv_bound_Ah) {
                for (int J = Jsq; J <= Jeq; ++J)
                    for (int I = Isq; I <= Ieq; ++I) A2(m->hv_Ah, I, J) = dmin(A2(m->hv_Ah, I, J), A2(m->hv_Ah_Max_xy, I, J));
            }
            for (int J = Jsq; J <= Jeq; ++J)
                for (int I = Isq; I <= Ieq; ++I) {
                    const double d_str = A2(m->hv_Ah, I, J) *
                        ((A2(m->hv_DY_dxBu, I, J) * ((A2(m->hv_Del2v, I + 1, J) * A2(m->IdyCv, I + 1, J)) - (A2(m->hv_Del2v, I, J) * A2(m->IdyCv, I, J)))) +
                         (A2(m->hv_DX_dyBu, I, J) * ((A2(m->hv_Del2u, I, J + 1) * A2(m->IdxCu, I, J + 1)) - (A2(m->hv_Del2u, I, J) * A2(m->IdxCu, I, J)))));
                    A2(m->hv_bhstr_xy, I, J) = d_str * (A2(m->hv_hq, I, J) * A2(m->hv_reduction_xy, I, J));
                    A2(m->hv_str_xy, I, J) = A2(m->hv_str_xy, I, J) + d_str;
                }
        }
        for (int J = Jsq; J <= Jeq; ++J)
            for (int I = Isq; I <= Ieq; ++I) {
                if (m->hv_no_slip)
                    A2(m->hv_str_xy, I, J) = A2(m->hv_str_xy, I, J) * (A2(m->hv_hq, I, J) * A2(m->hv_reduction_xy, I, J));
                else
                    A2(m->hv_str_xy, I, J) = A2(m->hv_str_xy, I, J) * (A2(m->hv_hq, I, J) * A2(m->mask2dBu, I, J) * A2(m->hv_reduction_xy, I, J));
            }
        for (int j = js; j <= je; ++j)
            for (int I = Isq; I <= Ieq; ++I)
                A3(diffu, I, j, k) = ((A2(m->IdxCu, I, j) * ((A2(m->hv_dx2q, I, j - 1) * A2(m->hv_str_xy, I, j - 1)) - (A2(m->hv_dx2q, I, j) * A2(m->hv_str_xy, I, j))) +
                                       A2(m->IdyCu, I, j) * ((A2(m->hv_dy2h, I, j) * A2(m->hv_str_xx, I, j)) - (A2(m->hv_dy2h, I + 1, j) * A2(m->hv_str_xx, I + 1, j)))) *
                                      A2(m->IareaCu, I, j)) / (A2(m->hv_h_u, I, j) + h_neglect);
        for (int J = Jsq; J <= Jeq; ++J)
            for (int i = is; i <= ie; ++i)
                A3(diffv, i, J, k) = ((A2(m->IdyCv, i, J) * ((A2(m->hv_dy2q, i - 1, J) * A2(m->hv_str_xy, i - 1, J)) - (A2(m->hv_dy2q, i, J) * A2(m->hv_str_xy, i, J))) -
                                       A2(m->IdxCv, i, J) * ((A2(m->hv_dx2h, i, J) * A2(m->hv_str_xx, i, J)) - (A2(m->hv_dx2h, i, J + 1) * A2(m->hv_str_xx, i, J + 1)))) *
                                      A2(m->IareaCv, i, J)) / (A2(m->hv_h_v, i, J) + h_neglect);
    }
}

/* --------------------------------------------------------------- driver --- */
/* compute_transports, rk2_serial_driver.F90:544-561 */
void om_compute_transports(om_model *m, const double *u, const double *v, const double *h, double *uh, double *vh)
{
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsd; j <= m->jed; ++j)
            for (int i = m->isd; i <= m->ied; ++i) {
                const int ip = i + 1 < m->ied ? i + 1 : m->ied, jp = j + 1 < m->jed ? j + 1 : m->jed;
                A3(uh, i, j, k) = A3(u, i, j, k) * 0.5 * (A3(h, i, j, k) + A3(h, ip, j, k)) * A2(m->dyCu, i, j);
                A3(vh, i, j, k) = A3(v, i, j, k) * 0.5 * (A3(h, i, j, k) + A3(h, i, jp, k)) * A2(m->dxCv, i, j);
            }
}

/* VelUpdate (:257-270 with fac=1, :337-350 with fac=0.5): out = base + fac*dt*(CA+diff) */
static void vel_update(om_model *m, double *ou, double *ov, const double *bu, const double *bv, double fdt)
{
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec - 1; ++i)
                A3(ou, i, j, k) = A3(bu, i, j, k) + fdt * (A3(m->CAu, i, j, k) + A3(m->diffu, i, j, k));
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec - 1; ++j)
            for (int i = m->isc; i <= m->iec; ++i)
                A3(ov, i, j, k) = A3(bv, i, j, k) + fdt * (A3(m->CAv, i, j, k) + A3(m->diffv, i, j, k));
}

void om_prepare(om_model *m)
{
    barotropic_init(m);
    if (periodic(m)) {
        halo_fill(m, m->u, m->nk); halo_fill(m, m->v, m->nk);
        halo_fill(m, m->h, m->nk); halo_fill(m, m->h0, m->nk);
        halo_fill(m, m->eta, 1); halo_fill(m, m->ubt, 1); halo_fill(m, m->vbt, 1);
    }
}


/* The meridional pass after the zonal one: its input is h after the zonal
 * pass with refilled periodic halos; its output keeps that halo. */
static void meridional_pass(om_model *m, const double *v, double dt)
{
    double *hx = (double *)malloc(m->N3 * sizeof(double));
    memcpy(hx, m->h, m->N3 * sizeof(double));
    if (periodic(m)) halo_fill(m, hx, m->nk);
    memcpy(m->h, hx, m->N3 * sizeof(double));
    om_continuity_ppm_y(m, v, hx, m->h, m->vh, dt);
    free(hx);
}

/* ------------------------------------------------------------ diagnostics */
/* post_data_2d / post_data_3d statistics (mom6_diag.F90:166-274): min, max,
 * mean, rms over the compute domain, summed in (k, j, i) order */
static void diag_stats(om_model *m, const double *f, int nk, double *out)
{
    double fmin = 1.79769313486231571e308, fmax = -1.79769313486231571e308, fmean = 0.0, frms = 0.0;
    long cnt = 0;
    for (int k = 1; k <= nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec; ++i) {
                const double x = A3(f, i, j, k);
                fmin = dmin(fmin, x);
                fmax = dmax(fmax, x);
                fmean = fmean + x;
                frms = frms + x * x;
                cnt = cnt + 1;
            }
    if (cnt > 0) {
        fmean = fmean / (double)cnt;
        frms = sqrt(frms / (double)cnt);
    }
    out[0] = fmin; out[1] = fmax; out[2] = fmean; out[3] = frms;
}

/* post_product_sum_u (mom6_diag.F90:276-308, u range) and post_product_sum_v
 * (:311-340, v range) into the one shared work array, then its statistics */
static void diag_product_sum(om_model *m, const double *a, const double *b, int v_range, double *out)
{
    const int ie = v_range ? m->iec : m->iec - 1, je = v_range ? m->jec - 1 : m->jec;
    for (int j = m->jsc; j <= je; ++j)
        for (int i = m->isc; i <= ie; ++i) A2(m->diag_work2d, i, j) = 0.0;
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= je; ++j)
            for (int i = m->isc; i <= ie; ++i)
                A2(m->diag_work2d, i, j) = A2(m->diag_work2d, i, j) + A3(a, i, j, k) * A3(b, i, j, k);
    diag_stats(m, m->diag_work2d, 1, out);
}

void om_diag(om_model *m, double *out)
{
    /* compute_and_post_KE (rk2_serial_driver.F90:664-693) */
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec; ++i) {
                const double u0 = A3(m->u, i, j, k), u1 = A3(m->u, i - 1, j, k);
                const double v0 = A3(m->v, i, j, k), v1 = A3(m->v, i, j - 1, k);
                A3(m->diag_ke, i, j, k) = 0.5 * (0.5 * (u0 * u0 + u1 * u1) + 0.5 * (v0 * v0 + v1 * v1));
            }
    diag_stats(m, m->diag_ke, m->nk, out);
    memcpy(out + 4, m->diag_posted, 8 * sizeof(double));
    /* compute_and_post_mass (:695-723) */
    for (size_t q = 0; q < m->N2; ++q) m->diag_mass[q] = 0.0;
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec; ++i) A2(m->diag_mass, i, j) = A2(m->diag_mass, i, j) + A3(m->h, i, j, k);
    diag_stats(m, m->diag_mass, 1, out + 12);
}

/* The coupled form's targets (no reference counterpart: the reference driver
 * passes uhbt = 0): the barotropic transports of the time-averaged barotropic
 * velocities of the btstep call just made, uhbt = Datu*ubt_av and
 * vhbt = Datv*vbt_av over the data domain (zero where btstep leaves the
 * averages zero: outside its narrow ranges, so the domain stays closed). */
static void coupled_targets(om_model *m)
{
    for (size_t q = 0; q < m->N2; ++q) {
        m->uhbt[q] = m->bt_Datu[q] * m->ubt_av[q];
        m->vhbt[q] = m->bt_Datv[q] * m->vbt_av[q];
    }
}

void om_step(om_model *m, int niter)
{
    const double dt = m->c.dt;
    const int per = periodic(m), nk = m->nk;
    for (int iter = 1; iter <= niter; ++iter) {
        memcpy(m->h, m->h0, m->N3 * sizeof(double));                       /* :222-228 */
        /* predictor */
        om_compute_transports(m, m->u, m->v, m->h, m->uh, m->vh);
        if (per) { halo_fill(m, m->uh, nk); halo_fill(m, m->vh, nk); }
        om_hor_visc(m, m->u, m->v, m->h, m->diffu, m->diffv);
        om_coradcalc(m, m->u, m->v, m->h, m->uh, m->vh, m->CAu, m->CAv);
        vel_update(m, m->up, m->vp, m->u, m->v, dt);
        om_vert_visc(m, m->up, m->vp, m->h, dt);
        om_btstep(m, m->eta, m->ubt, m->vbt, m->ubt_av, m->vbt_av, m->eta_av);
        if (m->c.coupled) coupled_targets(m);
        if (per) { halo_fill(m, m->up, nk); halo_fill(m, m->vp, nk); }
        om_continuity_ppm(m, m->up, m->h0, m->h, m->uh, dt);
        if (m->c.meridional) meridional_pass(m, m->vp, dt);
        if (m->c.diag) {   /* the predictor's posts (rk2_serial_driver.F90:300-308) */
            diag_product_sum(m, m->diffu, m->h, 0, m->diag_posted);
            diag_product_sum(m, m->diffv, m->h, 1, m->diag_posted + 4);
        }
        if (per) halo_fill(m, m->h, nk);
        /* corrector */
        om_compute_transports(m, m->up, m->vp, m->h, m->uh, m->vh);
        if (per) { halo_fill(m, m->uh, nk); halo_fill(m, m->vh, nk); }
        om_hor_visc(m, m->up, m->vp, m->h, m->diffu, m->diffv);
        om_coradcalc(m, m->up, m->vp, m->h, m->uh, m->vh, m->CAu, m->CAv);
        vel_update(m, m->u, m->v, m->u, m->v, 0.5 * dt);
        om_vert_visc(m, m->u, m->v, m->h, 0.5 * dt);
        om_btstep(m, m->eta_av, m->ubt_av, m->vbt_av, m->ubt_av, m->vbt_av, m->eta);
        if (m->c.coupled) coupled_targets(m);
        /* htmp <- h over the full data domain (the driver's loop bound
         * `i=isd,jed` at :375 is read as isd..ied; identical when ni == nj) */
        memcpy(m->htmp, m->h, m->N3 * sizeof(double));
        om_continuity_ppm(m, m->u, m->htmp, m->h, m->uh, 0.5 * dt);
        if (m->c.meridional) meridional_pass(m, m->v, 0.5 * dt);
        if (per) {
            halo_fill(m, m->u, nk); halo_fill(m, m->v, nk); halo_fill(m, m->h, nk);
            halo_fill(m, m->eta, 1); halo_fill(m, m->ubt, 1); halo_fill(m, m->vbt, 1);
        }
    }
}

/* verify_state, rk2_serial_driver.F90:563-623 */
void om_verify(om_model *m, double *out)
{
    double mass_init = 0, mass_final = 0, max_u = 0, max_v = 0, max_eta = 0, ke = 0;
    for (int k = 1; k <= m->nk; ++k)
        for (int j = m->jsc; j <= m->jec; ++j)
            for (int i = m->isc; i <= m->iec; ++i) {
                mass_init = mass_init + A3(m->h0, i, j, k) * A2(m->areaT, i, j);
                mass_final = mass_final + A3(m->h, i, j, k) * A2(m->areaT, i, j);
                max_u = dmax(max_u, fabs(A3(m->u, i, j, k)));
                max_v = dmax(max_v, fabs(A3(m->v, i, j, k)));
                const double uu = A3(m->u, i, j, k), vv = A3(m->v, i, j, k);
                ke = ke + 0.5 * (uu * uu + vv * vv) * A3(m->h, i, j, k) * A2(m->areaT, i, j);
            }
    for (int j = m->jsc; j <= m->jec; ++j)
        for (int i = m->isc; i <= m->iec; ++i) max_eta = dmax(max_eta, fabs(A2(m->eta, i, j)));
    out[0] = mass_init; out[1] = mass_final; out[2] = fabs(mass_final - mass_init) / mass_init;
    out[3] = max_u; out[4] = max_v; out[5] = max_eta; out[6] = ke;
}

/* ----------------------------------------------------------- lifecycle --- */
om_model *om_create(const om_config *c)
{
    om_model *m = (om_model *)calloc(1, sizeof(om_model));
    m->c = *c;
    m->ni = c->ni; m->nj = c->nj; m->nk = c->nk;
    m->isd = ISD; m->ied = c->ni + 2; m->jsd = JSD; m->jed = c->nj + 2;
    m->isc = m->isd + 3; m->iec = m->ied - 3; m->jsc = m->jsd + 3; m->jec = m->jed - 3;
    m->n1 = m->ied - m->isd + 1; m->n2 = m->jed - m->jsd + 1;
    m->N2 = (size_t)m->n1 * m->n2; m->N3 = m->N2 * (size_t)m->nk;
    m->first_direction = 0;
    const size_t N2 = m->N2, N3 = m->N3, N3p = m->N2 * (size_t)(m->nk + 1);
#define R2D(f) reg(m, #f, &m->f, N2, 0.0)
#define R3D(f) reg(m, #f, &m->f, N3, 0.0)
    R2D(IareaT); R2D(areaT); R2D(dxT); R2D(dyT); R2D(IdxT); R2D(IdyT); R2D(dxCu); R2D(dyCu);
    R2D(dy_Cu); R2D(IdxCu); R2D(IdyCu); R2D(dxCv); R2D(dyCv); R2D(IdxCv); R2D(IdyCv);
    R2D(IareaBu); R2D(areaBu); R2D(CoriolisBu); R2D(IareaCu); R2D(IareaCv); R2D(dxBu);
    R2D(dyBu); R2D(IdxBu); R2D(IdyBu); R2D(bathyT); R2D(mask2dT); R2D(mask2dBu);
    R2D(mask2dCu); R2D(mask2dCv);
    R3D(u); R3D(v); R3D(h); R3D(h0); R3D(htmp); R3D(uh); R3D(vh); R3D(CAu); R3D(CAv);
    R3D(up); R3D(vp); R3D(diffu); R3D(diffv);
    R2D(eta); R2D(ubt); R2D(vbt); R2D(ubt_av); R2D(vbt_av); R2D(eta_av);
    R2D(uhbt); R2D(vhbt); R3D(u_cor); R2D(du_cor);
    reg(m, "por_face_areaU", &m->por_face_areaU, N3, 1.0);
    reg(m, "visc_rem_u", &m->visc_rem_u, N3, 1.0);
    R2D(taux); R2D(tauy); R3D(Ray_u); R3D(Ray_v);
    R2D(diag_work2d); R3D(diag_ke); R2D(diag_mass);
    R2D(FA_u_EE); R2D(FA_u_E0); R2D(FA_u_W0); R2D(FA_u_WW); R2D(uBT_WW); R2D(uBT_EE);
    reg(m, "BT_h_u", &m->BT_h_u, N3, 0.0);
    R3D(v_cor); R2D(dv_cor);
    R2D(FA_v_NN); R2D(FA_v_N0); R2D(FA_v_S0); R2D(FA_v_SS); R2D(vBT_SS); R2D(vBT_NN);
    reg(m, "BT_h_v", &m->BT_h_v, N3, 0.0);
    R3D(cont_hW); R3D(cont_hE);
    R2D(cor_dvdx); R2D(cor_dudy); R2D(cor_rel_vort); R2D(cor_abs_vort); R2D(cor_q); R2D(cor_Ih_q);
    R2D(cor_hArea_u); R2D(cor_hArea_v); R2D(cor_Area_q); R2D(cor_KE); R2D(cor_a); R2D(cor_b);
    R2D(cor_c); R2D(cor_d);
    R2D(bt_eta); R2D(bt_eta_pred); R2D(bt_ubt); R2D(bt_vbt); R2D(bt_ubt_prev); R2D(bt_vbt_prev);
    R2D(bt_uhbt); R2D(bt_vhbt); R2D(bt_PFu); R2D(bt_PFv); R2D(bt_Cor_u); R2D(bt_Cor_v);
    R2D(bt_Datu); R2D(bt_Datv); R2D(bt_gtot_E); R2D(bt_gtot_W); R2D(bt_gtot_N); R2D(bt_gtot_S);
    reg(m, "bt_f_4_u", &m->bt_f_4_u, 4 * N2, 0.0);
    reg(m, "bt_f_4_v", &m->bt_f_4_v, 4 * N2, 0.0);
    R2D(bt_rem_u); R2D(bt_rem_v); R2D(bt_ubt_av); R2D(bt_vbt_av);
    reg(m, "vv_a_u", &m->vv_a_u, N3p, 0.0);
    reg(m, "vv_a_v", &m->vv_a_v, N3p, 0.0);
    R3D(vv_h_u); R3D(vv_h_v); R2D(vv_taux_bot); R2D(vv_tauy_bot);
    R2D(hv_Kh_bg_xx); R2D(hv_Kh_bg_xy); R2D(hv_Ah_bg_xx); R2D(hv_Ah_bg_xy);
    R2D(hv_Kh_Max_xx); R2D(hv_Kh_Max_xy); R2D(hv_Ah_Max_xx); R2D(hv_Ah_Max_xy);
    R2D(hv_Laplac2_const_xx); R2D(hv_Laplac2_const_xy); R2D(hv_Biharm_const_xx); R2D(hv_Biharm_const_xy);
    R2D(hv_dx2h); R2D(hv_dy2h); R2D(hv_dx2q); R2D(hv_dy2q); R2D(hv_DX_dyT); R2D(hv_DY_dxT);
    R2D(hv_DX_dyBu); R2D(hv_DY_dxBu); R2D(hv_Idx2dyCu); R2D(hv_Idxdy2u); R2D(hv_Idx2dyCv);
    R2D(hv_Idxdy2v); R2D(hv_reduction_xx); R2D(hv_reduction_xy);
    R2D(hv_dudx); R2D(hv_dvdy); R2D(hv_dvdx); R2D(hv_dudy); R2D(hv_sh_xx); R2D(hv_sh_xy);
    R2D(hv_str_xx); R2D(hv_str_xy); R2D(hv_bhstr_xx); R2D(hv_bhstr_xy); R2D(hv_h_u); R2D(hv_h_v);
    R2D(hv_hq); R2D(hv_Kh); R2D(hv_Ah); R2D(hv_Shear_mag); R2D(hv_Del2u); R2D(hv_Del2v);
    R2D(hv_hrat_min); R2D(hv_visc_bound_rem);
    R2D(hv_Laplac3_const_xx); R2D(hv_Laplac3_const_xy); R2D(hv_Biharm6_const_xx); R2D(hv_Biharm6_const_xy);
    R2D(hv_vort_xy); R2D(hv_vort_xy_dx); R2D(hv_vort_xy_dy); R2D(hv_grad_vort_mag_h); R2D(hv_grad_vort_mag_q);
    R2D(hv_vert_vort_mag); R2D(hv_Del2vort_q);
#undef R2D
#undef R3D
    grid_init(m);
    /* continuity_init (:53-78) */
    m->upwind_1st = c->upwind_1st; m->simple_2nd = c->simple_2nd; m->monotonic = c->monotonic;
    m->tol_eta = 1.0e-12; m->tol_vel = 3.0e8; m->CFL_limit_adjust = 0.5;
    m->aggress_adjust = 0; m->vol_CFL = 0; m->better_iter = 1; m->use_visc_rem_max = 1;
    m->marginal_faces = 1;
    coriolis_init(m);
    /* vert_visc_init (:60-112) */
    m->vv_Kv = c->Kv; m->vv_Kv_ml = c->Kv_ml; m->vv_Kv_extra_bbl = c->Kv_extra_bbl;
    m->vv_Hmix = c->Hmix; m->vv_Hbbl = c->Hbbl;
    hor_visc_init(m);
    /* initialize_forces_visc Rayleigh part (rk2_serial_driver.F90:524-540);
     * taux/tauy are inputs set by the caller */
    for (int j = m->jsd; j <= m->jed; ++j)
        for (int i = m->isd; i <= m->ied; ++i) {
            A3(m->Ray_u, i, j, m->nk) = c->ray_bottom;
            A3(m->Ray_v, i, j, m->nk) = c->ray_bottom;
        }
    barotropic_init(m);
    return m;
}

void om_destroy(om_model *m)
{
    if (!m) return;
    om_destroy(m->mt);
    for (int q = 0; q < m->ntab; ++q) free(*m->tab[q].p);
    free(m->tab);
    free(m);
}

double *om_field(om_model *m, const char *name, long *n)
{
    for (int q = 0; q < m->ntab; ++q)
        if (strcmp(m->tab[q].name, name) == 0) {
            if (n) *n = m->tab[q].n;
            return *m->tab[q].p;
        }
    if (n) *n = 0;
    return NULL;
}

int om_dims(om_model *m, int *o)
{
    o[0] = m->isd; o[1] = m->ied; o[2] = m->jsd; o[3] = m->jed;
    o[4] = m->isc; o[5] = m->iec; o[6] = m->jsc; o[7] = m->jec;
    return m->nk;
}
