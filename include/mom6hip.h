/*
 * mom6hip.h -- C ABI of the gfx950 (MI355X) split-RK2 dynamical core.
 *
 * Drop-in replacement for the reference's src/cuda_c layer.  Each module keeps
 * the control-structure pattern of the cuda_c Fortran wrappers (init copies the
 * 2D metrics to the device, compute takes device pointers, end frees), so a
 * Fortran iso_c_binding module can expose the same subroutines
 * (continuity_init_cuda_c -> mom6h_continuity_init, ...).  The kernels follow
 * the SERIAL reference semantics (the src/serial modules), not the cuda_c ones.
 *
 * Conventions
 *  - Arrays: double, Fortran column-major over the data domain isd..ied x
 *    jsd..jed (x 1..nk for 3D).  Bounds are 1-based and inclusive.
 *  - Pointers named d_* (and every array argument of a compute call) are device
 *    pointers; init calls take host arrays (h_*) and copy them.
 *  - `stream` is a hipStream_t (NULL = default stream).  Compute calls are
 *    asynchronous on it; they never synchronise the device.
 *  - Every function returns 0 (MOM6H_OK) or an error code; the last error text
 *    is available from mom6h_last_error().  The reference's launch_* entry
 *    points return nothing and never check launches; these do.
 */
#ifndef MOM6HIP_H
#define MOM6HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOM6H_OK 0
#define MOM6H_EINVAL 1001      /* bad argument (bounds, NULL, unknown name) */
#define MOM6H_ENOMEM 1002      /* device allocation failed */
#define MOM6H_ESTATE 1003      /* call out of order (e.g. step before prepare) */

const char *mom6h_last_error(void);
const char *mom6h_version(void);

/* ---------------------------------------------------------------- runtime --
 * Replace cuda_{set_device,device_synchronize,malloc,free,memcpy_*,memset}_c
 * (src/cuda_c/cuda_helpers.h:10-17, cuda_helpers.cu:8-48). */
int mom6h_set_device(int device);
int mom6h_device_synchronize(void);
int mom6h_malloc(void **d_ptr, size_t bytes);
int mom6h_free(void *d_ptr);
int mom6h_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int mom6h_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int mom6h_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
int mom6h_memset(void *d_ptr, int value, size_t bytes);
int mom6h_stream_create(void **stream);           /* default priority */
int mom6h_stream_destroy(void *stream);
int mom6h_stream_synchronize(void *stream);

/* ------------------------------------------------------------ continuity --
 * continuity_init_cuda_c (src/cuda_c/mom6_continuity_cuda_c.F90:269-340),
 * continuity_PPM_cuda_c (:444-617), continuity_end_cuda_c (:345-384);
 * arithmetic of src/serial/mom6_continuity_serial.F90:100-1352.
 * dxT, dxCu and mask2dCu are required here (optional in cuda_c). */
typedef struct mom6h_continuity_cs mom6h_continuity_cs;
typedef struct mom6h_bt_cont {
    /* BT_cont_type (src/common/mom6_types.F90:89-110): device arrays */
    double *FA_u_EE, *FA_u_E0, *FA_u_W0, *FA_u_WW, *uBT_WW, *uBT_EE; /* 2D */
    double *h_u;                                                       /* 3D */
} mom6h_bt_cont;

int mom6h_continuity_init(mom6h_continuity_cs **cs, int isd, int ied, int jsd, int jed,
                          int isc, int iec, int jsc, int jec, int nk,
                          const double *h_IareaT, const double *h_IdxT, const double *h_dy_Cu,
                          const double *h_mask2dT, int monotonic, const double *h_dxT,
                          const double *h_dxCu, const double *h_mask2dCu);
/* u, hin -> h (compute domain), uh; uhbt (2D, NULL = zero transport target),
 * bt_cont (NULL = skip), u_cor (3D) and du_cor (2D) optional (NULL = not stored).
 * When hin_halo_to_h is non-zero, h outside the compute domain is set to hin
 * (the driver's h reset fused into the call); otherwise it is left untouched.
 * hin and h are distinct arrays (the stencils read hin while h is written). */
int mom6h_continuity_ppm(mom6h_continuity_cs *cs, const double *u, const double *hin, double *h,
                         double *uh, double dt, const double *uhbt, const mom6h_bt_cont *bt_cont,
                         double *u_cor, double *du_cor, int hin_halo_to_h, void *stream);
int mom6h_continuity_end(mom6h_continuity_cs *cs);
/* Meridional pass.  No counterpart in the reference, whose continuity is
 * zonal only (continuity_PPM_cuda_c, src/serial/mom6_continuity_serial.F90:
 * 100-145): MOM6's meridional_mass_flux, i.e. the zonal algorithm with
 * (i, u, IdxT, dy_Cu, dxT, dxCu, mask2dCu) exchanged for (j, v, IdyT, dx_Cv,
 * dyT, dyCv, mask2dCv).  set_meridional copies those y metrics (host arrays);
 * ppm_y then takes v, hin -> h (compute domain), vh on the faces (i, J),
 * i = isc..iec, J = jsc-1..jec, with the same optional outputs as the zonal
 * call; bt_cont receives the v-side BT_cont fields in the same slots
 * (FA_u_EE -> FA_v_NN, FA_u_E0 -> FA_v_N0, FA_u_W0 -> FA_v_S0, FA_u_WW ->
 * FA_v_SS, uBT_WW -> vBT_SS, uBT_EE -> vBT_NN, h_u -> h_v).  Needs nk <= 80. */
int mom6h_continuity_set_meridional(mom6h_continuity_cs *cs, const double *h_IdyT, const double *h_dx_Cv,
                                    const double *h_dyT, const double *h_dyCv, const double *h_mask2dCv);
int mom6h_continuity_ppm_y(mom6h_continuity_cs *cs, const double *v, const double *hin, double *h,
                           double *vh, double dt, const double *vhbt, const mom6h_bt_cont *bt_cont,
                           double *v_cor, double *dv_cor, int hin_halo_to_h, void *stream);
/* alloc_BT_cont_type_cuda_c / dealloc (:389-435): zeroed device arrays */
int mom6h_bt_cont_alloc(mom6h_bt_cont *bt, int isd, int ied, int jsd, int jed, int nk);
int mom6h_bt_cont_free(mom6h_bt_cont *bt);

/* -------------------------------------------------------------- coriolis --
 * coriolis_init_cuda_c (src/cuda_c/mom6_coriolis_cuda_c.F90:164-227),
 * CorAdCalc_cuda_c (:270-356); arithmetic of
 * src/serial/mom6_coriolis_serial.F90:112-252.  scheme: 1 Sadourny75 energy,
 * 2 Arakawa-Hsu90, 3 Arakawa-Lamb81. */
typedef struct mom6h_coriolis_cs mom6h_coriolis_cs;
int mom6h_coriolis_init(mom6h_coriolis_cs **cs, int isd, int ied, int jsd, int jed,
                        int isc, int iec, int jsc, int jec, int nk,
                        const double *h_areaT, const double *h_IareaBu, const double *h_CoriolisBu,
                        const double *h_mask2dBu, const double *h_dyCv, const double *h_dxCu,
                        const double *h_dyCu, const double *h_dxCv, const double *h_IdxCu,
                        const double *h_IdyCv, int scheme);
int mom6h_coradcalc(mom6h_coriolis_cs *cs, const double *u, const double *v, const double *h,
                    const double *uh, const double *vh, double *CAu, double *CAv, void *stream);
int mom6h_coriolis_end(mom6h_coriolis_cs *cs);

/* ------------------------------------------------------------ barotropic --
 * barotropic_init_cuda_c (src/cuda_c/mom6_barotropic_cuda_c.F90:226-300),
 * btstep_cuda_c (:305-433), barotropic_end_cuda_c (:629-669); arithmetic of
 * src/serial/mom6_barotropic_serial.F90:178-356.  Each substep is ONE launch
 * of an LDS-tiled kernel (eta predictor, pressure forces, both velocity
 * updates, transports, free surface, time averages on a tile plus its halo)
 * instead of ~9 dependent launches; grids whose coefficients are uniform take
 * a variant that holds them as values.  periodic_x/periodic_y select the
 * split-API ranges (btstep_do_step, :379-474) with a halo refill of
 * ubt, vbt, eta after every substep but the last. */
typedef struct mom6h_barotropic_cs mom6h_barotropic_cs;
int mom6h_barotropic_init(mom6h_barotropic_cs **cs, int isd, int ied, int jsd, int jed,
                          int isc, int iec, int jsc, int jec, int nstep, double dt, double bebt,
                          int first_direction, const double *h_Datu, const double *h_Datv,
                          const double *h_gtot_E, const double *h_gtot_W, const double *h_gtot_N,
                          const double *h_gtot_S, const double *h_f_4_u, const double *h_f_4_v,
                          const double *h_bt_rem_u, const double *h_bt_rem_v,
                          const double *h_IareaT, const double *h_IdxCu, const double *h_IdyCv,
                          int periodic_x, int periodic_y);
/* eta_in/ubt_in/vbt_in may alias u_av/v_av/eta_av (the driver's corrector call). */
int mom6h_btstep(mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in,
                 const double *vbt_in, double *u_av, double *v_av, double *eta_av, void *stream);
/* mom6h_btstep without device allocation or a control structure's scratch:
 * no counterpart in the reference.  For callers that cannot allocate (code
 * under stream/graph capture, serving loops, harnesses that forbid
 * hipMalloc): the two state sets and the accumulators live in a static
 * device array of 64 MiB (8 planes) per GPU, so the grid may have at most
 * 1,048,576 points n1*n2 (MOM6H_EINVAL beyond).  The array is shared by
 * every call on a GPU and cleared on each call: calls that overlap on one GPU
 * (different streams) race.  Same results as mom6h_btstep. */
int mom6h_btstep_noalloc(mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in,
                         const double *vbt_in, double *u_av, double *v_av, double *eta_av, void *stream);
int mom6h_barotropic_end(mom6h_barotropic_cs *cs);

/* ------------------------------------------------------------- vert_visc --
 * vert_visc_init_cuda_c (src/cuda_c/mom6_vert_visc_cuda_c.F90:89-125),
 * vert_visc_cra_cuda_c (:130-188); arithmetic of
 * src/serial/mom6_vert_visc_serial.F90:138-287 (coef) and :394-506 (apply),
 * including the Rayleigh drag the cuda_c path lacks.  ray_u_bot/ray_v_bot are
 * 2D Rayleigh coefficients of the bottom layer (NULL = none). */
typedef struct mom6h_vert_visc_cs mom6h_vert_visc_cs;
int mom6h_vert_visc_init(mom6h_vert_visc_cs **cs, int isd, int ied, int jsd, int jed,
                         int isc, int iec, int jsc, int jec, int nk, const double *h_mask2dCu,
                         const double *h_mask2dCv, double Kv, double Kv_ml, double Kv_extra_bbl,
                         double Hmix, double Hbbl);
int mom6h_vert_visc(mom6h_vert_visc_cs *cs, double *u, double *v, const double *h, double dt,
                    const double *taux, const double *tauy, const double *ray_u_bot,
                    const double *ray_v_bot, void *stream);
/* mom6h_vert_visc with the reference's general Rayleigh drag: ray_u/ray_v
 * are 3D (n1*n2*nk, vertvisc_type Ray_u/Ray_v, src/common/mom6_types.F90:120-125),
 * Ray(k) enters every level of the tridiagonal system and the bottom stress
 * adds sum_k rho0*Ray(k)*u(k) (vert_visc_apply :427-458). */
int mom6h_vert_visc_ray3d(mom6h_vert_visc_cs *cs, double *u, double *v, const double *h, double dt,
                          const double *taux, const double *tauy, const double *ray_u, const double *ray_v,
                          void *stream);
/* Whether mom6h_vert_visc stores the coupling coefficients and effective
 * thicknesses (default 1).  With 0 they stay on chip: used when nothing
 * reads them after the call. */
int mom6h_vert_visc_keep_coefficients(mom6h_vert_visc_cs *cs, int keep);
/* device pointers of the coefficient arrays (a_u/a_v: nk+1 levels) and bottom stress */
int mom6h_vert_visc_arrays(mom6h_vert_visc_cs *cs, double **a_u, double **a_v, double **h_u,
                           double **h_v, double **taux_bot, double **tauy_bot);
int mom6h_vert_visc_end(mom6h_vert_visc_cs *cs);

/* -------------------------------------------------------------- hor_visc --
 * hor_visc_init_cuda_c (src/cuda_c/mom6_hor_visc_cuda_c.F90:117-190),
 * hor_visc_cuda_c (:237-298); arithmetic of
 * src/serial/mom6_hor_visc_serial.F90:464-1109 (Laplacian, biharmonic). */
typedef struct mom6h_hor_visc_cs mom6h_hor_visc_cs;
int mom6h_hor_visc_init(mom6h_hor_visc_cs **cs, int isd, int ied, int jsd, int jed,
                        int isc, int iec, int jsc, int jec, int nz, double Kh, double h_neglect,
                        const double *h_DY_dxT, const double *h_DX_dyT, const double *h_DY_dxBu,
                        const double *h_DX_dyBu, const double *h_IdyCu, const double *h_IdxCu,
                        const double *h_IdyCv, const double *h_IdxCv, const double *h_IareaCu,
                        const double *h_IareaCv, const double *h_mask2dT, const double *h_mask2dBu,
                        const double *h_reduction_xx, const double *h_reduction_xy,
                        const double *h_dy2h, const double *h_dx2h, const double *h_dy2q,
                        const double *h_dx2q);
/* Enables the biharmonic term (hor_visc_init biharmonic=.true., Ah; :160-286)
 * and selects free-slip (0) or no-slip (1) boundaries; laplacian=0 drops Kh. */
int mom6h_hor_visc_set_options(mom6h_hor_visc_cs *cs, int laplacian, int biharmonic, double Ah,
                               int no_slip, const double *h_Idx2dyCu, const double *h_Idxdy2u,
                               const double *h_Idx2dyCv, const double *h_Idxdy2v);
/* hor_visc_init's Smagorinsky=, Leith= and better_bound= options
 * (mom6_hor_visc_serial.F90:160-286; Smagorinsky_Kh/_Ah, Leith_Kh/_Ah follow
 * the Laplacian/biharmonic switches; better_bound sets the Kh and Ah bounds)
 * with the grid metrics they need: areaT (the Smagorinsky, Leith and bound
 * constants), dxBu, dyBu, IdxBu, IdyBu (Leith).  Before the first call. */
int mom6h_hor_visc_set_closures(mom6h_hor_visc_cs *cs, int smagorinsky, int leith, int better_bound,
                                const double *h_areaT, const double *h_dxBu, const double *h_dyBu,
                                const double *h_IdxBu, const double *h_IdyBu);
int mom6h_hor_visc(mom6h_hor_visc_cs *cs, const double *u, const double *v, const double *h,
                   double *diffu, double *diffv, void *stream);
int mom6h_hor_visc_end(mom6h_hor_visc_cs *cs);

/* ------------------------------------------------------- driver kernels --
 * The rk2 drivers' inline loops (app/rk2_serial_driver.F90). */
/* compute_transports (:544-561), full data domain */
int mom6h_transports(int isd, int ied, int jsd, int jed, int nk, const double *u, const double *v,
                     const double *h, const double *dyCu, const double *dxCv, double *uh,
                     double *vh, void *stream);
/* VelUpdate (:257-270, :337-350): ou = bu + fac_dt*(CAu+diffu) on the
 * reference ranges (ou may alias bu). */
int mom6h_vel_update(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                     double fac_dt, const double *bu, const double *bv, const double *CAu,
                     const double *diffu, const double *CAv, const double *diffv, double *ou,
                     double *ov, void *stream);
/* Periodic refill of the 3-wide halo (the on-device replacement of
 * halo_exchange_3d/2d, src/common/mom6_mpi_halo.F90) */
int mom6h_halo_fill(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                    int periodic_x, int periodic_y, double *a, void *stream);

/* Barotropic CFL of the split time step (check_bt_cfl,
 * app/rk2_serial_driver.F90:625-662): 1D and 2D-diagonal CFL of the gravity
 * wave sqrt(g*depth) at dt/nsteps, the substep count the reference recommends
 * (ceiling(2.5*dt*c/dx)), status 0 = OK, 1 = CFL_2d > 0.7, 2 = CFL_2d > 0.9.
 * Host only. */
int mom6h_bt_cfl(double dt, int nsteps, double dx_km, double depth, double *cfl, double *cfl_2d,
                 int *min_nsteps, int *status);

/* Seeded initial state on the device (no counterpart in the reference, whose
 * driver uses an analytic state): N(mean, sigma) samples from SplitMix64 and
 * Irwin-Hall sums of 12 uniforms, bit-identical to the host generator.
 * add_mean = 0 omits the mean; mask2d (n2 values, NULL = none) multiplies
 * every layer. */
int mom6h_seeded_normal(double *d_out, long count, unsigned long long seed, int field, int add_mean,
                        double mean, double sigma, const double *d_mask2d, long n2, void *stream);

/* ----------------------------------------------------------- diagnostics --
 * The serial driver's diagnostics (diag_mode = 1) on the device:
 * post_product_sum_u/v (src/common/mom6_diag.F90:276-340), the KE and mass
 * fields of compute_and_post_KE/_mass (app/rk2_serial_driver.F90:664-723) and
 * the min/max/mean/rms statistics of post_data_2d/3d (mom6_diag.F90:166-274),
 * which the caller prints or writes (the reference's binary dump layout is
 * written on the host).  Per-point fields are bit-identical to the serial
 * loops; mean and rms come from a deterministic double-double reduction (the
 * serial sum differs by its own rounding). */
/* work_2d = sum_k a*b on the u range (position 0: isc..iec-1 x jsc..jec) or
 * the v range (1: isc..iec x jsc..jec-1); other points are left as they are */
int mom6h_diag_product_sum(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                           const double *a, const double *b, int position, double *work_2d, void *stream);
/* KE at the compute h points (3D) */
int mom6h_diag_ke(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk, const double *u,
                  const double *v, double *ke, void *stream);
/* mass = column sum of h on the compute points, zero elsewhere (2D) */
int mom6h_diag_mass(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk, const double *h,
                    double *mass, void *stream);
/* doubles of device workspace mom6h_diag_stats needs */
long mom6h_diag_stats_work_doubles(void);
/* d_out[0..3] = min, max, mean, rms of field over the compute domain and
 * nk layers (nk = 1: a 2D field) */
int mom6h_diag_stats(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk,
                     const double *field, double *d_work, double *d_out, void *stream);

/* ------------------------------------------------------------ the model --
 * A whole run of app/rk2_serial_driver.F90 on one GPU: grid, control
 * structures and state on the device; mom6h_model_step advances niter RK2
 * steps (:218-399) with every module on the GPU and the barotropic chain on a
 * second stream.  Fields are addressed by the driver's names: u v h h0 htmp
 * uh vh CAu CAv up vp diffu diffv eta ubt vbt ubt_av vbt_av eta_av u_cor du_cor
 * taux tauy ray_u ray_v CoriolisBu mask2dT mask2dCu mask2dCv mask2dBu, and the
 * BT_cont fields FA_u_EE FA_u_E0 FA_u_W0 FA_u_WW uBT_WW uBT_EE BT_h_u. */
typedef struct mom6h_config {
    int ni, nj, nk, bt_nsteps;
    double dt, dx_km;
    int periodic_x, periodic_y, walls_y;
    double Kh, Ah;
    int laplacian, biharmonic, no_slip;
    double Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl;
    int coriolis_scheme, monotonic;
    int meridional;   /* continuity also takes the meridional pass (mom6h_continuity_ppm_y) */
    int precision;    /* 0: FP64 throughout (the reference's arithmetic); 1: mixed -- FP32 layer
                         fields (u v h h0 htmp up vp uh vh u_cor BT_h_u ...) and FP32 stencil
                         arithmetic (PPM continuity with Kahan-compensated column sums, Coriolis/KE,
                         Laplacian viscosity); FP64 vertical-viscosity solve, barotropic solver,
                         2D fields and BT_cont 2D fields.  Needs the fused stencils: uniform
                         metrics, Sadourny Coriolis, Laplacian viscosity, nk <= 80. */
    int smagorinsky, leith, better_bound;   /* hor_visc closures (mom6h_hor_visc_set_closures);
                                               FP64 form, module-by-module stencils */
    int coupled;      /* continuity's barotropic-transport targets from the barotropic solver:
                         uhbt = Datu*ubt_av (and vhbt = Datv*vbt_av for the meridional pass) of the
                         btstep call before each continuity call; the reference passes uhbt = 0.
                         FP64 form; the barotropic chain then runs in the 3D stream's order. */
    int ray3d;        /* Rayleigh drag from the 3D fields "ray_u3d"/"ray_v3d" (mom6h_vert_visc_ray3d)
                         instead of the 2D bottom-layer "ray_u"/"ray_v"; FP64 form */
} mom6h_config;

typedef struct mom6h_model mom6h_model;
int mom6h_model_create(mom6h_model **m, const mom6h_config *c);
/* Device address and element count of a field; its element type is FP64,
 * or FP32 for the layer fields of the mixed-precision form
 * (mom6h_model_field_bytes gives 8 or 4). */
int mom6h_model_field(mom6h_model *m, const char *name, void **d_ptr, long *count);
int mom6h_model_field_bytes(mom6h_model *m, const char *name, int *elem_bytes);
/* Host FP64 values in and out (converted for FP32 fields), and a fill from
 * FP64 values already on the device (d_src, count elements). */
int mom6h_model_upload(mom6h_model *m, const char *name, const double *h_src, long count);
int mom6h_model_download(mom6h_model *m, const char *name, double *h_dst, long count);
int mom6h_model_fill(mom6h_model *m, const char *name, const double *d_src, long count);
/* after the inputs are uploaded: barotropic coefficients from CoriolisBu,
 * wall masks, periodic fill of the initial state */
int mom6h_model_prepare(mom6h_model *m);
int mom6h_model_step(mom6h_model *m, int niter);
int mom6h_model_synchronize(mom6h_model *m);
/* Options: "graph" (1 = capture one step as two hipGraphs, the 3D chain on
 * the model's first stream and the barotropic chain on its second, and replay
 * them per step, joined once per call as the eager path is; default 0), "momentum_tile" (rows of the fused momentum
 * tile: 4, 8 or 16; default 8), "momentum_splits" (blocks per tile sharing its
 * layers, 1..nk; default 1), "bt_tile" (rows of the barotropic substep tile,
 * 4, 8 or 16; after prepare; default 4), "bt_substeps" (barotropic substeps per
 * launch between the first and the last of a call: 1 = one launch per substep,
 * up to 5 = the blocked kernel; after prepare; default 5), "bt_rows" (1 = the
 * blocked kernel stages Datv and the Coriolis factor per row when both are
 * constant along the compute rows, 0 = always as 2D fields; after prepare;
 * default 1), "fused" (1 = fused momentum stencil and
 * on-chip vert_visc coefficients where the grid allows, default 1; 0 = the
 * module-by-module path), "profile" (1 = record per-stage HIP events; the
 * graph is not used while profiling), "serialize" (1 = run the barotropic
 * chain on the 3D stream, one step after the other, so that per-stage events
 * time each kernel without the overlap; default 0). */
int mom6h_model_set_option(mom6h_model *m, const char *key, double value);
/* Per-stage GPU time (ms) accumulated while "profile" is on; names are
 * written comma-separated into names (size nbytes). Returns the stage count. */
int mom6h_model_stage_times(mom6h_model *m, double *ms, int max_stages, char *names, int nbytes);
int mom6h_model_main_stream(mom6h_model *m, void **stream);
/* Times the fused momentum stencil (reps launches of the predictor's, whose
 * outputs every step rewrites) for each launch shape -- tile rows 4, 8, 16 x
 * layer splits 1, 2, 4, 8 -- keeps the fastest as options "momentum_tile" and
 * "momentum_splits" and writes the timings into report. */
int mom6h_model_autotune(mom6h_model *m, int reps, char *report, int nbytes);
/* With option "diag" = 1 (FP64 form) every step posts the predictor's
 * diffu_sum and diffv_sum statistics; this call adds KE and mass of the
 * current state and returns the 16 values KE, diffu_sum, diffv_sum, mass
 * (min, max, mean, rms each): the serial driver's Diag lines
 * (rk2_serial_driver.F90:152-170, 300-308, 386-394).  Synchronous. */
int mom6h_model_diagnostics(mom6h_model *m, double *out16);
int mom6h_model_destroy(mom6h_model *m);

/* ---------------------------------------------------------- the ensemble --
 * Several models (members: own grid size, nk and state) advanced together:
 * each stage of the step is one launch over the tiles of every member, the
 * barotropic chains of all members share a second stream.  Members must share
 * bt_nsteps, precision, meridional and biharmonic, and need the fused
 * stencils (uniform metrics, Sadourny Coriolis, Laplacian viscosity).  Fill a
 * member's fields through the model API on the handle mom6h_ensemble_member
 * returns (upload / fill / download), then prepare once.  A member's results
 * equal those of the same model stepped alone.  No counterpart in the
 * reference (its drivers run one grid). */
typedef struct mom6h_ensemble mom6h_ensemble;
int mom6h_ensemble_create(mom6h_ensemble **e, const mom6h_config *member_cfg, int n_members);
int mom6h_ensemble_member(mom6h_ensemble *e, int index, mom6h_model **m);
int mom6h_ensemble_prepare(mom6h_ensemble *e);
int mom6h_ensemble_step(mom6h_ensemble *e, int niter);
int mom6h_ensemble_synchronize(mom6h_ensemble *e);
/* Options: "graph" (1 = capture one step, replay it per step), "momentum_tile"
 * (rows of the fused momentum tile: 4, 8 or 16; default 8), "momentum_splits"
 * (blocks per tile sharing its layers; after prepare; default 1), "bt_tile" (4, 8 or 16),
 * "bt_substeps" (barotropic substeps per blocked launch, 1-5; default 5). */
int mom6h_ensemble_set_option(mom6h_ensemble *e, const char *key, double value);
/* Times the fused momentum launch of the ensemble (reps launches) for every
 * launch shape (tile rows x layer splits), keeps the fastest and writes the
 * timings into report.  It runs
 * only the predictor's momentum stencil, whose outputs every step rewrites,
 * so the state is unchanged. */
int mom6h_ensemble_autotune(mom6h_ensemble *e, int reps, char *report, int nbytes);
int mom6h_ensemble_destroy(mom6h_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif /* MOM6HIP_H */
