// mom6h_internal.h -- shared definitions of the gfx950 split-RK2 library.
//
// Layout (same as the Fortran callers): column-major, i fastest, then j, then
// k; 1-based Fortran indices over the data domain isd..ied x jsd..jed.  Device
// code indexes with the 1-based values through IX2/IX3 so every loop bound reads
// like the Fortran it mirrors.  All arithmetic is FP64 and the library is built
// with -ffp-contract=off: no FMA contraction, IEEE division, Fortran's
// left-to-right association kept by explicit parentheses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/mom6hip.h"

namespace mom6h {

// Bounds of one grid, 1-based and inclusive, plus derived strides.
struct Dims {
    int isd, ied, jsd, jed, isc, iec, jsc, jec, nk;
    int n1, n2;          // ied-isd+1, jed-jsd+1
    long long s2;        // n1*n2 : stride of k
};

inline Dims make_dims(int isd, int ied, int jsd, int jed, int isc, int iec, int jsc, int jec, int nk)
{
    Dims d;
    d.isd = isd; d.ied = ied; d.jsd = jsd; d.jed = jed;
    d.isc = isc; d.iec = iec; d.jsc = jsc; d.jec = jec; d.nk = nk;
    d.n1 = ied - isd + 1; d.n2 = jed - jsd + 1;
    d.s2 = (long long)d.n1 * d.n2;
    return d;
}

__host__ __device__ inline long long IX2(const Dims &d, int i, int j)
{
    return (long long)(j - d.jsd) * d.n1 + (i - d.isd);
}
__host__ __device__ inline long long IX3(const Dims &d, int i, int j, int k)
{
    return ((long long)(k - 1) * d.n2 + (j - d.jsd)) * d.n1 + (i - d.isd);
}

// Block-to-tile mapping for the 8 XCDs of MI355X.  The hardware hands
// consecutive workgroups to the XCDs in turn, each XCD with its own L2; this
// renumbers them so that XCD x works on one contiguous range of tiles, and
// tiles that share halos (and 2D fields read by every layer) meet in the same
// L2.  A bijection on [0, nb) for any nb.
__device__ inline int xcd_tile(int b, int nb)
{
    const int m = nb / 8, r = nb % 8;
    const int x = b % 8, q = b / 8;
    return x < r ? x * (m + 1) + q : r * (m + 1) + (x - r) * m + q;
}

__device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__device__ inline double dmin(double a, double b) { return a < b ? a : b; }

// Halo modes of a grid (periodic refill of the 3-wide halo, or none).
struct Halo {
    int px, py;   // periodic in x / y
};

// ------------------------------------------------------------------ errors --
int set_error(int code, const char *fmt, ...);
#define MOM6H_CHECK(expr)                                                          \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::mom6h::set_error((int)e_, "%s:%d %s: %s", __FILE__, __LINE__, \
                                      #expr, hipGetErrorString(e_));               \
    } while (0)
#define MOM6H_LAUNCH_CHECK() MOM6H_CHECK(hipGetLastError())
#define MOM6H_REQUIRE(cond, ...)                                                   \
    do {                                                                           \
        if (!(cond)) return ::mom6h::set_error(MOM6H_EINVAL, __VA_ARGS__);         \
    } while (0)

// Device array helpers (synchronous; set-up time only).
int dev_alloc(double **p, size_t count, bool zero);
int dev_upload(double **p, const double *host, size_t count);   // alloc + copy
void dev_free(double *p);

// Halo fill of nk layers (nk = 1 for 2D) on a stream; no-op when not periodic.
int halo_fill(const Dims &d, const Halo &hm, double *a, int nk, hipStream_t s);
int halo_fill(const Dims &d, const Halo &hm, float *a, int nk, hipStream_t s);

// Element-wise FP64 <-> FP32 copies on a stream (round to nearest / exact)
int narrow_to_f32(const double *src, float *dst, long long n, hipStream_t s);
int widen_to_f64(const float *src, double *dst, long long n, hipStream_t s);
int dev_alloc_bytes(void **p, size_t bytes, bool zero);

}  // namespace mom6h

struct mom6h_vert_visc_cs;
struct mom6h_continuity_cs;

namespace mom6h {

// The mixed-precision forms (FP32 layer fields) of the column physics.
int vert_visc_f32(mom6h_vert_visc_cs *cs, float *u, float *v, const float *h, double dt, const double *taux,
                  const double *tauy, const double *ray_u_bot, const double *ray_v_bot, hipStream_t stream);
// continuity_PPM on FP32 thickness and velocity: the zonal pass (with the
// convergence) and the meridional pass; BT_cont 2D fields and du_cor FP64,
// the marginal thickness bt_h_u FP32.
int continuity_f32(mom6h_continuity_cs *cs, const float *u, const float *hin, float *h, float *uh, double dt,
                   const mom6h_bt_cont *bt, float *bt_h_u, float *u_cor, double *du_cor, int hin_halo_to_h,
                   hipStream_t s);
int continuity_y_f32(mom6h_continuity_cs *cs, const float *v, const float *hin, float *h, float *vh, double dt,
                     const mom6h_bt_cont *bt, float *bt_h_v, float *v_cor, double *dv_cor, int hin_halo_to_h,
                     hipStream_t s);

}  // namespace mom6h

// ------------------------------------------------------------------------
// Control structures (opaque in the C ABI).
// ------------------------------------------------------------------------
struct mom6h_continuity_cs {
    mom6h::Dims d;
    int monotonic, upwind_1st, simple_2nd;
    double tol_eta, tol_vel, CFL_limit_adjust;
    double *IareaT, *IdxT, *dy_Cu, *mask2dT, *dxT, *dxCu, *mask2dCu;
    double *hW, *hE;         // PPM edge values (3D scratch of the module kernels)
    double *IdyT, *dx_Cv, *dyT, *dyCv, *mask2dCv;   // meridional pass (NULL until set)
    long long max_lds;       // device limit on LDS per block
};

struct mom6h_coriolis_cs {
    mom6h::Dims d;
    int scheme;
    double *areaT, *IareaBu, *CoriolisBu, *mask2dBu, *dyCv, *dxCu, *dyCu, *dxCv, *IdxCu, *IdyCv;
    double *Area_q;          // sum of the 4 surrounding cell areas
    double *q, *KE;          // 3D potential vorticity / kinetic energy scratch
};

struct mom6h_barotropic_cs {
    mom6h::Dims d;
    mom6h::Halo halo;
    int nstep, first_direction;
    double dtbt, bebt, dgeo_de;
    double *Datu, *Datv, *gtot_E, *gtot_W, *gtot_N, *gtot_S, *f_4_u, *f_4_v;
    double *bt_rem_u, *bt_rem_v, *IareaT, *IdxCu, *IdyCv;
    // uniform grids: the coefficients as values, f4 = f_4 per point (2D)
    int uniform;
    double u_Datu, u_gtot, u_rem_u, u_rem_v, u_IareaT, u_IdxCu, u_IdyCv;
    double *f4;
    // state (two sets, ping-pong across substeps) and time averages; zero at
    // init, as the reference's never-written cells
    double *eta, *ubt, *vbt, *eta_b, *ubt_b, *vbt_b, *ubt_av, *vbt_av;
};

struct mom6h_vert_visc_cs {
    mom6h::Dims d;
    double Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl;
    int keep_coefs;          // store a_u/a_v/h_u/h_v (the reference's CS arrays)
    long long max_lds;       // device limit on LDS per block
    double *mask2dCu, *mask2dCv;
    double *a_u, *a_v;       // coupling coefficients, nk+1 levels
    double *h_u, *h_v;       // effective thicknesses at velocity points
    double *c1;              // tridiagonal scratch (u and v columns)
    double *taux_bot, *tauy_bot;
};

struct mom6h_hor_visc_cs {
    mom6h::Dims d;
    double Kh, Ah, h_neglect;
    int laplacian, biharmonic, no_slip;
    int calls;               // number of completed calls (the first layer of the first call differs)
    double *DY_dxT, *DX_dyT, *DY_dxBu, *DX_dyBu, *IdyCu, *IdxCu, *IdyCv, *IdxCv;
    double *IareaCu, *IareaCv, *mask2dT, *mask2dBu, *reduction_xx, *reduction_xy;
    double *dy2h, *dx2h, *dy2q, *dx2q;
    double *Idx2dyCu, *Idxdy2u, *Idx2dyCv, *Idxdy2v;
    double *Del2u, *Del2v;   // biharmonic scratch (3D, zero outside their ranges)
};
