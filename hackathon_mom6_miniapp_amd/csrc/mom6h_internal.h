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

// true and the common value when the 2D host array a holds one value on
// ilo..ihi x jlo..jhi (clipped to the data domain)
inline bool uniform_over(const Dims &d, const double *a, int ilo, int ihi, int jlo, int jhi, double *v)
{
    ilo = ilo < d.isd ? d.isd : ilo; ihi = ihi > d.ied ? d.ied : ihi;
    jlo = jlo < d.jsd ? d.jsd : jlo; jhi = jhi > d.jed ? d.jed : jhi;
    const double a0 = a[IX2(d, ilo, jlo)];
    for (int j = jlo; j <= jhi; ++j)
        for (int i = ilo; i <= ihi; ++i)
            if (!(a[IX2(d, i, j)] == a0)) return false;
    *v = a0;
    return true;
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

// Several grids (the members of an ensemble) in one launch: the blocks of
// member m are first[m] .. first[m+1]-1, laid out as that member's own
// launch would be (nbx x nby tiles, x fastest, then any further dimension).
constexpr int kMaxMembers = 64;
struct BatchMap {
    int n;                          // members
    int first[kMaxMembers + 1];     // first block of each member; first[n] = total
    int nbx[kMaxMembers], nby[kMaxMembers];
};

// member of global block b and the block's index within that member
__device__ inline void batch_locate(const BatchMap &mp, int b, int &m, int &local)
{
    int q = 0;
    while (q + 1 < mp.n && b >= mp.first[q + 1]) ++q;
    m = q;
    local = b - mp.first[q];
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

// Halo fill of many arrays (the members of an ensemble) in one launch per
// direction; a job with px = py = 0 is skipped.
template <typename T>
struct HaloJob {
    Dims d;
    T *a;
    int nk, px, py;
};
template <typename T>
int halo_batch(const HaloJob<T> *d_jobs, const HaloJob<T> *h_jobs, int n, hipStream_t s);

// Element-wise FP64 <-> FP32 copies on a stream (round to nearest / exact)
int narrow_to_f32(const double *src, float *dst, long long n, hipStream_t s);
int widen_to_f64(const float *src, double *dst, long long n, hipStream_t s);
int dev_alloc_bytes(void **p, size_t bytes, bool zero);

// ------------------------------------------------- kernel argument blocks
// vert_visc of one grid.
// T: storage type of u, v and h (double, or float in the mixed-precision
// form); the coefficients and the tridiagonal solve are FP64 either way, and
// between sweeps B and C the forward-eliminated velocities live in the
// velocity array itself (rounded to T).
template <typename T>
struct VVArgs {
    Dims d;
    T *u, *v;
    const T *h;
    double dt;
    const double *taux, *tauy, *ray_u, *ray_v, *mask2dCu, *mask2dCv;
    double *a_u, *a_v, *h_u, *h_v, *taux_bot, *tauy_bot;
    double Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl;
    const double *ray_u3, *ray_v3;   // 3D Rayleigh coefficients (the 3D variant; NULL otherwise)
};

// Arguments of one pass.  Names are those of the zonal pass; the meridional
// pass passes v, vh, ... and its own metrics in the same slots:
//   IdxT -> IdyT, dy_Cu -> dx_Cv, dxT -> dyT, dxCu -> dyCv, mask2dCu -> mask2dCv.
// T is the type of the 3D layer fields and of the flux arithmetic: double
// (the reference's FP64) or float (the mixed-precision form); metrics, the
// 2D BT_cont outputs and du_cor stay FP64 in memory.
template <typename T>
struct ContArgsT {
    Dims d;
    const T *u, *hin;
    T *h, *uh;
    double dt;
    const double *uhbt;
    const double *IareaT, *IdxT, *dy_Cu, *dxT, *dxCu, *mask2dCu, *mask2dT;
    double tol_eta0, tol_vel, CFL_limit_adjust, h_min;
    T *u_cor;
    double *du_cor;
    double *FA_u_EE, *FA_u_E0, *FA_u_W0, *FA_u_WW, *uBT_WW, *uBT_EE;
    T *bt_h_u;
    int monotonic;   // read by the batched launch (the single one takes it as an argument)
};
using ContArgs = ContArgsT<double>;

// btstep: the scalars of one grid, its coefficients and one substep's buffers.
struct BT {
    Dims d;
    int nstep, first_direction, extended, px, py;
    double dtbt, bebt, dgeo_de;
    double *ubt_av, *vbt_av;   // running time averages, accumulated in place
};

// Metric access.  BTArrays reads every coefficient from its 2D array (the
// general grid); BTUniform holds the coefficients that are the same at every
// point as values and keeps only Datv (closed faces of walls) and the
// Coriolis factors (f_4 = one value per point, equal over the four
// components) as arrays.  The arithmetic is the same for both.
struct BTArrays {
    const double *Datu, *Datv, *gtot_E, *gtot_W, *gtot_N, *gtot_S, *f_4_u, *f_4_v;
    const double *bt_rem_u, *bt_rem_v, *IareaT, *IdxCu, *IdyCv;
    __device__ double datu(long long c) const { return Datu[c]; }
    __device__ double datv(long long c) const { return Datv[c]; }
    __device__ double gE(long long c) const { return gtot_E[c]; }
    __device__ double gW(long long c) const { return gtot_W[c]; }
    __device__ double gN(long long c) const { return gtot_N[c]; }
    __device__ double gS(long long c) const { return gtot_S[c]; }
    __device__ double f4u(int comp, long long c) const { return f_4_u[4 * c + (comp - 1)]; }
    __device__ double f4v(int comp, long long c) const { return f_4_v[4 * c + (comp - 1)]; }
    __device__ double remu(long long c) const { return bt_rem_u[c]; }
    __device__ double remv(long long c) const { return bt_rem_v[c]; }
    __device__ double iareaT(long long c) const { return IareaT[c]; }
    __device__ double idxCu(long long c) const { return IdxCu[c]; }
    __device__ double idyCv(long long c) const { return IdyCv[c]; }
};

struct BTUniform {
    double Datu, gtot, bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv;
    const double *Datv, *f4;
    // Datv and f4 per row (jsd..jed) when each is constant along the rows of
    // the compute domain, else NULL
    const double *Datv_row = nullptr, *f4_row = nullptr;
    __device__ double datu(long long) const { return Datu; }
    __device__ double datv(long long c) const { return Datv[c]; }
    __device__ double gE(long long) const { return gtot; }
    __device__ double gW(long long) const { return gtot; }
    __device__ double gN(long long) const { return gtot; }
    __device__ double gS(long long) const { return gtot; }
    __device__ double f4u(int, long long c) const { return f4[c]; }
    __device__ double f4v(int, long long c) const { return f4[c]; }
    __device__ double remu(long long) const { return bt_rem_u; }
    __device__ double remv(long long) const { return bt_rem_v; }
    __device__ double iareaT(long long) const { return IareaT; }
    __device__ double idxCu(long long) const { return IdxCu; }
    __device__ double idyCv(long long) const { return IdyCv; }
};

struct BTStep {
    const double *eta_o, *ubt_o, *vbt_o;   // state before the substep
    double *eta_n, *ubt_n, *vbt_n;         // state after it
    double *out_u, *out_v, *out_eta;       // written on the last substep only
    int n;                                 // substep number, 1-based
};


// One member of a batched btstep call: its grid and coefficients, its two
// state buffer sets, and the call's inputs and outputs.
struct BTMember {
    BT p;
    BTUniform m;
    int blocked_ok;             // the blocked kernel may run its interior substeps
    double *A[3], *B[3];        // eta, ubt, vbt: A after the copy-in, B the other set
    const double *in[3];        // eta_in, ubt_in, vbt_in
    double *out[3];             // u_av, v_av, eta_av
};
int btstep_batch(const BTMember *d_args, const BTMember *h_args, int n, hipStream_t s, int ty, int per_launch);

template <typename T>
int vert_visc_batch(const VVArgs<T> *d_args, const VVArgs<T> *h_args, int n, long long max_lds, hipStream_t s);
template <typename T>
int continuity_batch(const ContArgsT<T> *d_args, const ContArgsT<T> *h_args, int n, bool y_pass, bool halo_copy,
                     long long max_lds, hipStream_t s);

}  // namespace mom6h

struct mom6h_vert_visc_cs;
struct mom6h_continuity_cs;

struct mom6h_barotropic_cs;

namespace mom6h {
int bt_member(const mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in, const double *vbt_in,
              double *u_av, double *v_av, double *eta_av, BTMember *out);
template <typename T>
ContArgsT<T> continuity_args(const mom6h_continuity_cs *cs, bool y_pass, const T *u, const T *hin, T *h, T *uh,
                             double dt, const mom6h_bt_cont *bt, T *bt_h_u, T *u_cor, double *du_cor);
}  // namespace mom6h

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
    // every metric constant over the points the stencil reads: the values
    // (the tiled kernel), else the two general kernels
    int uniform = 0;
    int general = 0;         // use the general kernels even on a uniform grid
    double u_areaT = 0, u_IareaBu = 0, u_dyCv = 0, u_dxCu = 0, u_dyCu = 0, u_dxCv = 0, u_IdxCu = 0, u_IdyCv = 0;
    double u_Area_q = 0;
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
    int tile_rows;           // rows of the substep tile (4, 8 or 16)
    int substeps;            // substeps per launch of the blocked kernel (1: one launch per substep)
    int blocked_ok;          // uniform coefficients whose periodic Datv halo rows equal their images
    double *Datv_row, *f4_row;   // Datv and f4 per row when constant along the compute rows (else NULL)
    int use_rows;                // stage Datv and f4 per row when the grid allows (1, the default) or as 2D fields
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

namespace mom6h {
// hor_visc's metric values on a grid whose metric arrays are constant
struct HVUniform {
    double DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, IdxCv, IareaCu, IareaCv;
    double red_xx, red_xy, dy2h, dx2h, dy2q, dx2q;
    double Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v;
};
}  // namespace mom6h

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
    // Smagorinsky / Leith viscosities and the thickness-aware bounds
    // (hor_visc_init Smagorinsky=, Leith=, better_bound=)
    int smagorinsky = 0, leith = 0, better_bound = 0;
    double *areaT = nullptr, *dxBu = nullptr, *dyBu = nullptr, *IdxBu = nullptr, *IdyBu = nullptr;
    // q-point Kh and Ah of the last layer of the previous call, read by the
    // first layer's stress ring (zero before the first call)
    double *ring_Kh = nullptr, *ring_Ah = nullptr;
    // every metric constant over the points the stencil reads (is-2..ie+2 x
    // js-2..je+2): the values, for the tiled kernel; bih_uniform likewise for
    // the four biharmonic metrics
    int uniform = 0, bih_uniform = 0;
    int general = 0;         // use the general kernel even on a uniform grid
    mom6h::HVUniform g{};
};
