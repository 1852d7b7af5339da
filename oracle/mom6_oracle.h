/*
 * mom6_oracle.h -- CPU restatement of the MOM6 mini-app serial split-RK2 step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; nothing in the product (libmom6hip.so, the Fortran/Python host code)
 * links or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * It restates, loop by loop and in the same floating-point association, the
 * reference serial CPU path:
 *   src/common/mom6_types.F90                 (grid, constants)
 *   src/serial/mom6_continuity_serial.F90     (zonal PPM continuity)
 *   src/serial/mom6_coriolis_serial.F90       (CorAdCalc)
 *   src/serial/mom6_barotropic_serial.F90     (btstep + split API)
 *   src/serial/mom6_vert_visc_serial.F90      (vert_visc_coef / apply)
 *   src/serial/mom6_hor_visc_serial.F90       (hor_visc)
 *   app/rk2_serial_driver.F90                 (step order, transports, updates)
 * Memory the reference allocates without initialising and then reads is
 * zero here (work arrays are calloc'ed and persist across calls, exactly as the
 * reference's allocatable control-structure arrays do).
 *
 * Layout: Fortran column-major, 1-based (i fastest, then j, then k), data
 * domain isd=1..ied=ni+2, jsd=1..jed=nj+2, compute domain isc=4..iec=ni-1.
 */
#ifndef MOM6_ORACLE_H
#define MOM6_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary modes (bc_x, bc_y).  STATIC is the reference single-domain
 * behaviour: halos are never refilled.  PERIODIC fills the 3-wide halo from
 * the opposite side of the compute domain at the exchange points of
 * app/rk2_mpi_serial_driver.F90 (:191-197, 247-248, 301-303, 313-314, 329,
 * 343-344, 397-399, 425-430).  WALL (bc_y only) closes the domain at the
 * southern and northern compute rows with land masks. */
enum { OM_BC_STATIC = 0, OM_BC_PERIODIC = 1, OM_BC_WALL = 2 };

typedef struct om_config {
    int ni, nj, nk;          /* driver CLI sizes (data domain ni+2 x nj+2) */
    int bt_nsteps;
    double dt;               /* 300 s in the reference driver */
    double dx_km;            /* 10 km */
    int bc_x, bc_y;
    int coriolis_mode;       /* 0: reference beta-plane f0=2*Omega*sin45; 1: f0 + beta*y given below */
    double f0, beta;
    double Kh, Ah;           /* hor_visc background viscosities */
    int laplacian, biharmonic, smagorinsky, no_slip, bound_Kh, better_bound;
    double Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl;
    double ray_bottom;       /* Rayleigh drag in the bottom layer [m/s] */
    int coriolis_scheme;     /* 1 Sadourny75, 2 Arakawa-Hsu90, 3 Arakawa-Lamb81 */
    int monotonic;           /* continuity limiter: 0 positive-definite, 1 CW84 */
    int simple_2nd, upwind_1st;
    int meridional;          /* continuity also takes a meridional pass after the zonal one
                                (no reference counterpart; see om_continuity_ppm_y) */
    int diag;                /* the driver's diag_mode: post the predictor's diffu/diffv sums */
    int leith;               /* hor_visc_init Leith=.true. (Leith_Kh, and Leith_Ah with biharmonic) */
    int coupled;             /* continuity's barotropic-transport targets from btstep: uhbt = Datu*ubt_av,
                                vhbt = Datv*vbt_av after each btstep call (the reference leaves uhbt = 0) */
} om_config;

void om_config_default(om_config *c);

typedef struct om_model om_model;

om_model *om_create(const om_config *c);
void om_destroy(om_model *m);

/* Raw access to a named field (grid metric, state, work or CS array).
 * Returns NULL for unknown names; *n receives the element count. */
double *om_field(om_model *m, const char *name, long *n);
int om_dims(om_model *m, int *out8);   /* isd ied jsd jed isc iec jsc jec */

/* Fill the periodic halos of the initial state (no-op for static BCs). */
void om_prepare(om_model *m);
/* Advance niter RK2 steps (rk2_serial_driver.F90:218-399). */
void om_step(om_model *m, int niter);

/* Module-level entry points on explicit arrays, with the model's grid and
 * control structures (same argument meaning as the serial subroutines). */
void om_compute_transports(om_model *m, const double *u, const double *v, const double *h,
                           double *uh, double *vh);
void om_hor_visc(om_model *m, const double *u, const double *v, const double *h,
                 double *diffu, double *diffv);
void om_coradcalc(om_model *m, const double *u, const double *v, const double *h,
                  const double *uh, const double *vh, double *CAu, double *CAv);
void om_vert_visc(om_model *m, double *u, double *v, const double *h, double dt);
void om_btstep(om_model *m, const double *eta_in, const double *ubt_in, const double *vbt_in,
               double *u_av, double *v_av, double *eta_av);
void om_continuity_ppm(om_model *m, const double *u, const double *hin, double *h, double *uh,
                       double dt);
/* Meridional PPM pass (no reference counterpart): the zonal pass on the
 * transposed grid and state; v, hin -> h (compute domain), vh; v_cor, dv_cor
 * and the BT_cont v-side fields (FA_v_*, vBT_*, BT_h_v) are stored in the model. */
void om_continuity_ppm_y(om_model *m, const double *v, const double *hin, double *h, double *vh,
                         double dt);
/* periodic refill of the 3-wide halo of a (nk layers; no-op for static BCs) */
void om_halo_fill(om_model *m, double *a, int nk);
/* verify_state (rk2_serial_driver.F90:563-623): out[7] = mass_init, mass_final,
 * rel_error, max_u, max_v, max_eta, ke_total. */
void om_verify(om_model *m, double *out7);
/* The driver's diagnostics with diag_mode = 1 (rk2_serial_driver.F90:152-170,
 * 300-308, 386-394; src/common/mom6_diag.F90): out[16] = min, max, mean, rms of
 * KE (of the current state), diffu_sum, diffv_sum (the last predictor's
 * posts) and mass (of the current state).  The fields are kept as
 * "diag_ke", "diag_work2d", "diag_mass". */
void om_diag(om_model *m, double *out16);

#ifdef __cplusplus
}
#endif
#endif
