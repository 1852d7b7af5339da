// mom6h_continuity.hip -- zonal PPM continuity on gfx950.
//
// Arithmetic of src/serial/mom6_continuity_serial.F90 (the serial oracle, not
// the cuda_c kernels): positive-definite limiter by default (monotonic=.false.,
// :66-76), REAL(4) tolerance literals (:826-830, :856-866, :993-994, :1086,
// :1105) and REAL(4) limiter locals (:1266-1268, :1311-1315), Newton
// bookkeeping that keeps the last evaluated layer fluxes (:877-914).
//
// The driver passes visc_rem_u = por_face_areaU = 1 everywhere
// (mom6_continuity_serial.F90:63-64).  Multiplying or dividing by 1.0 is exact
// in IEEE arithmetic, so the kernels drop those factors and stay bit-identical.
//
// Kernels
//   k_ppm_edges     (i,j,k) threads  -> h_W, h_E          zonal_edge_thickness
//   k_mass_flux     one thread per u-column (I,j), k loops: layer fluxes, CFL
//                   bracket, Newton to the barotropic transport, BT_cont and
//                   h_u, u_cor/du_cor                       zonal_mass_flux
//   k_convergence   (i,j,k) threads  -> h                  continuity_zonal_convergence
#include "mom6h_internal.h"

namespace mom6h {

// ------------------------------------------------------------ PPM edges ---
// PPM_reconstruction_x (:1124-1247) slope at one point.
__device__ inline double ppm_slope(const Dims &d, const double *__restrict__ h,
                                   const double *__restrict__ mT, int i, int j, int k)
{
    if ((mT[IX2(d, i - 1, j)] * mT[IX2(d, i, j)] * mT[IX2(d, i + 1, j)]) == 0.0) return 0.0;
    const double hp = h[IX3(d, i + 1, j, k)], hm = h[IX3(d, i - 1, j, k)], h0 = h[IX3(d, i, j, k)];
    double s = 0.5 * (hp - hm);
    const double dMx = dmax(dmax(hp, hm), h0) - h0;
    const double dMn = h0 - dmin(dmin(hp, hm), h0);
    return copysign(1.0, s) * dmin(fabs(s), 2.0 * dmin(dMx, dMn));
}

// Edge values of cell (i, j, k) (zonal_edge_thickness :185-207 with
// PPM_reconstruction_x :1124-1247 and its limiter).
__device__ inline void ppm_edge(const Dims &d, const double *__restrict__ h, const double *__restrict__ mT,
                                int i, int j, int k, int upwind_1st, int simple_2nd, int monotonic,
                                double h_min, double &hW, double &hE)
{
    const double hi = h[IX3(d, i, j, k)];
    if (upwind_1st) {  // (:199-207): no reconstruction, no limiter
        hW = hi;
        hE = hi;
        return;
    }
    const double mm = mT[IX2(d, i - 1, j)], mp = mT[IX2(d, i + 1, j)];
    const double h_im1 = mm * h[IX3(d, i - 1, j, k)] + (1.0 - mm) * hi;
    const double h_ip1 = mp * h[IX3(d, i + 1, j, k)] + (1.0 - mp) * hi;
    double hl, hr;
    if (simple_2nd) {
        hl = 0.5 * (h_im1 + hi);
        hr = 0.5 * (h_ip1 + hi);
    } else {
        const double oneSixth = 1.0 / 6.0;
        const double sm = ppm_slope(d, h, mT, i - 1, j, k);
        const double s0 = ppm_slope(d, h, mT, i, j, k);
        const double sp = ppm_slope(d, h, mT, i + 1, j, k);
        hl = 0.5 * (h_im1 + hi) + oneSixth * (sm - s0);
        hr = 0.5 * (h_ip1 + hi) + oneSixth * (s0 - sp);
    }
    if (monotonic) {  // PPM_limit_CW84 (:1298-1338), REAL(4) locals
        const float h_i = (float)hi;
        if ((hr - (double)h_i) * ((double)h_i - hl) <= 0.0) {
            hl = (double)h_i;
            hr = (double)h_i;
        } else {
            const float RLdiff = (float)(hr - hl);
            const float RLmean = (float)(0.5 * (hr + hl));
            const float FunFac = 6.0f * RLdiff * (h_i - RLmean);
            const float RLdiff2 = RLdiff * RLdiff;
            if (FunFac > RLdiff2) hl = (double)(3.0f * h_i) - 2.0 * hr;
            if (FunFac < -RLdiff2) hr = (double)(3.0f * h_i) - 2.0 * hl;
        }
    } else {  // PPM_limit_pos (:1253-1294), REAL(4) curv, dh, scale
        const float curv = (float)(3.0 * ((hl + hr) - 2.0 * hi));
        if ((double)curv > 0.0) {
            const float dh = (float)(hr - hl);
            if (fabsf(dh) < curv) {
                if (hi <= h_min) {
                    hl = hi;
                    hr = hi;
                } else {
                    const double lhs = (double)(12.0f * curv) * (hi - h_min);
                    const float rhs = curv * curv + 3.0f * (dh * dh);
                    if (lhs < (double)rhs) {
                        const float scale = (float)(lhs / (double)rhs);
                        hl = hi + (double)scale * (hl - hi);
                        hr = hi + (double)scale * (hr - hi);
                    }
                }
            }
        }
    }
    hW = hl;
    hE = hr;
}

__global__ void k_ppm_edges(Dims d, const double *__restrict__ h, const double *__restrict__ mT,
                            double *__restrict__ hW, double *__restrict__ hE, int upwind_1st,
                            int simple_2nd, int monotonic, double h_min)
{
    const int i = d.isc - 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.iec + 1) return;
    const long long c = IX3(d, i, j, k);
    double hl, hr;
    ppm_edge(d, h, mT, i, j, k, upwind_1st, simple_2nd, monotonic, h_min, hl, hr);
    hW[c] = hl;
    hE[c] = hr;
}

// ------------------------------------------------------- layer flux ---
// zonal_flux_layer (:516-598) for one face with visc_rem = por = 1:
// uh = (dy*u)*(...), h_marg; duhdu = dy*h_marg.
__device__ inline void zflux(double u, double dt, double dy, double IdxT0, double IdxT1, double hW0,
                             double hE0, double h0, double hW1, double hE1, double h1, double &uh,
                             double &hmarg)
{
    if (u > 0.0) {
        const double CFL = u * dt * IdxT0;
        const double curv_3 = (hW0 + hE0) - 2.0 * h0;
        uh = (dy * u) * (hE0 + CFL * (0.5 * (hW0 - hE0) + curv_3 * (CFL - 1.5)));
        hmarg = hE0 + CFL * ((hW0 - hE0) + 3.0 * curv_3 * (CFL - 1.0));
    } else if (u < 0.0) {
        const double CFL = -(u * dt * IdxT1);
        const double curv_3 = (hW1 + hE1) - 2.0 * h1;
        uh = (dy * u) * (hW1 + CFL * (0.5 * (hE1 - hW1) + curv_3 * (CFL - 1.5)));
        hmarg = hW1 + CFL * ((hE1 - hW1) + 3.0 * curv_3 * (CFL - 1.0));
    } else {
        uh = 0.0;
        hmarg = 0.5 * (hW1 + hE0);
    }
}

// Per-column view of the six thickness values a flux at face I needs.
struct Column {
    const double *u, *hW, *hE, *h;  // base pointers
    long long ci, ci1;              // offset of cell i and i+1 at k=1
    long long cu;                   // offset of face I at k=1
    long long s2;                   // k stride
    double dt, dy, IdxT0, IdxT1;
    __device__ inline void eval(int k, double un, double &uh, double &hm) const
    {
        const long long o = (long long)(k - 1) * s2;
        zflux(un, dt, dy, IdxT0, IdxT1, hW[ci + o], hE[ci + o], h[ci + o], hW[ci1 + o], hE[ci1 + o],
              h[ci1 + o], uh, hm);
    }
    __device__ inline double uk(int k) const { return u[cu + (long long)(k - 1) * s2]; }
};

// zonal_flux_adjust (:739-916) for one column.  uh3 != nullptr: uh_3d present
// (layer fluxes written back, evaluated also at itt = max_itts).
__device__ double flux_adjust(const Column &col, int nk, double uhbt, double uh_tot_0,
                              double duhdu_tot_0, double du_max_CFL, double du_min_CFL,
                              double dtIarea, double tol_eta0, double tol_vel, double *uh3)
{
    const int max_itts = 20;
    double du = 0.0, du_max = du_max_CFL, du_min = du_min_CFL;
    double uh_err = uh_tot_0 - uhbt, duhdu_tot = duhdu_tot_0;
    double uh_err_best = fabs(uh_err);
    bool do_I = true;
    for (int itt = 1; itt <= max_itts; ++itt) {
        double tol_eta;
        if (itt <= 1) tol_eta = (double)1e-6f * tol_eta0;
        else if (itt == 2) tol_eta = (double)1e-4f * tol_eta0;
        else if (itt == 3) tol_eta = (double)1e-2f * tol_eta0;
        else tol_eta = tol_eta0;
        if (uh_err > 0.0) du_max = du;
        else if (uh_err < 0.0) du_min = du;
        else do_I = false;
        if (do_I) {
            const double aerr = fabs(uh_err);
            if ((dtIarea * aerr > tol_eta) || ((aerr > tol_vel * duhdu_tot) || (aerr > uh_err_best))) {
                const double ddu = -uh_err / duhdu_tot;
                const double du_prev = du;
                du = du + ddu;
                if (fabs(ddu) < (double)1.0e-15f * fabs(du)) {
                    do_I = false;
                } else if (ddu > 0.0) {
                    if (du >= du_max) {
                        du = 0.5 * (du_prev + du_max);
                        if (du_max - du_prev < (double)1.0e-15f * fabs(du)) do_I = false;
                    }
                } else {
                    if (du <= du_min) {
                        du = 0.5 * (du_prev + du_min);
                        if (du_prev - du_min < (double)1.0e-15f * fabs(du)) do_I = false;
                    }
                }
            } else {
                do_I = false;
            }
        }
        // A column that is done leaves its du untouched for the rest of the
        // row's iterations in the reference, so it can stop here.
        if (!do_I) break;
        if ((itt < max_itts) || uh3) {
            double se = -uhbt, sd = 0.0;
            for (int k = 1; k <= nk; ++k) {
                double uh, hm;
                col.eval(k, col.uk(k) + du, uh, hm);
                if (uh3) uh3[col.cu + (long long)(k - 1) * col.s2] = uh;
                se = se + uh;
                sd = sd + col.dy * hm;
            }
            if (itt < max_itts) {
                uh_err = se;
                duhdu_tot = sd;
                uh_err_best = dmin(uh_err_best, fabs(uh_err));
            }
        }
    }
    return du;
}

__global__ void k_mass_flux(Dims d, const double *__restrict__ u, const double *__restrict__ hin,
                            const double *__restrict__ hW, const double *__restrict__ hE,
                            double *__restrict__ uh, double dt, const double *__restrict__ uhbt,
                            const double *__restrict__ IareaT, const double *__restrict__ IdxT,
                            const double *__restrict__ dy_Cu, const double *__restrict__ dxT,
                            const double *__restrict__ dxCu, const double *__restrict__ mask2dCu,
                            double tol_eta0, double tol_vel, double CFL_limit_adjust,
                            double *__restrict__ u_cor, double *__restrict__ du_cor,
                            double *__restrict__ FA_u_EE, double *__restrict__ FA_u_E0,
                            double *__restrict__ FA_u_W0, double *__restrict__ FA_u_WW,
                            double *__restrict__ uBT_WW, double *__restrict__ uBT_EE,
                            double *__restrict__ bt_h_u, int set_bt_cont)
{
    const int I = d.isc - 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc + blockIdx.y;
    if (I > d.iec) return;
    const int i = I;
    const int nk = d.nk;
    Column col;
    col.u = u; col.hW = hW; col.hE = hE; col.h = hin;
    col.ci = IX3(d, i, j, 1); col.ci1 = IX3(d, i + 1, j, 1); col.cu = IX3(d, I, j, 1);
    col.s2 = d.s2; col.dt = dt;
    const long long c2 = IX2(d, I, j);
    col.dy = dy_Cu[c2];
    col.IdxT0 = IdxT[IX2(d, i, j)];
    col.IdxT1 = IdxT[IX2(d, i + 1, j)];
    const double mCu = mask2dCu[c2];

    // layer fluxes and their sums (zonal_mass_flux :300-355)
    const double CFL_dt = CFL_limit_adjust / dt;
    const double dx_W = dxT[IX2(d, i, j)], dx_E = dxT[IX2(d, i + 1, j)];
    double du_max_CFL = 2.0 * (CFL_dt * dx_W) * 1.0;     // I_vrm = 1/visc_rem_max = 1
    double du_min_CFL = -(2.0 * (CFL_dt * dx_E) * 1.0);
    double uh_tot_0 = 0.0, duhdu_tot_0 = 0.0;
    for (int k = 1; k <= nk; ++k) {
        const double uu = col.uk(k);
        double f, hm;
        col.eval(k, uu, f, hm);
        uh[col.cu + (long long)(k - 1) * d.s2] = f;
        duhdu_tot_0 = duhdu_tot_0 + col.dy * hm;
        uh_tot_0 = uh_tot_0 + f;
        // CFL bracket, use_visc_rem branch without aggress_adjust (:378-393)
        if (du_max_CFL > dx_W * CFL_dt - uu * mCu) du_max_CFL = (dx_W * CFL_dt - uu);
        if (du_min_CFL < -dx_E * CFL_dt - uu * mCu) du_min_CFL = -(dx_E * CFL_dt + uu);
    }
    du_max_CFL = dmax(du_max_CFL, 0.0);
    du_min_CFL = dmin(du_min_CFL, 0.0);

    const double dtIarea = dt * dmin(IareaT[IX2(d, i, j)], IareaT[IX2(d, i + 1, j)]);
    const double bt = uhbt ? uhbt[c2] : 0.0;
    const double du = flux_adjust(col, nk, bt, uh_tot_0, duhdu_tot_0, du_max_CFL, du_min_CFL,
                                  dtIarea, tol_eta0, tol_vel, uh);
    if (du_cor) du_cor[c2] = du;
    if (u_cor)
        for (int k = 1; k <= nk; ++k) u_cor[col.cu + (long long)(k - 1) * d.s2] = col.uk(k) + du;
    if (!set_bt_cont) return;

    // set_zonal_BT_cont (:920-1121); the zero-transport correction equals du
    // whenever the target transport is zero (identical Newton iterations).
    const double du0 = (bt == 0.0) ? du
                                   : flux_adjust(col, nk, 0.0, uh_tot_0, duhdu_tot_0, du_max_CFL,
                                                 du_min_CFL, dtIarea, tol_eta0, tol_vel, nullptr);
    const double Idt = 1.0 / dt;
    const double CFL_min = (double)1e-6f;
    const double du_CFL = (CFL_min * Idt) * dxCu[c2];
    double duR = dmin(0.0, du0 - du_CFL);
    double duL = dmax(0.0, du0 + du_CFL);
    for (int k = 1; k <= nk; ++k) {   // visc_rem_lim = max(1, 0.1) = 1
        const double uu = col.uk(k);
        if (uu + duR > -du_CFL) duR = -(uu + du_CFL);
        if (uu + duL < du_CFL) duL = -(uu - du_CFL);
    }
    double FAmt_0 = 0.0, FAmt_L = 0.0, FAmt_R = 0.0, uhtot_L = 0.0, uhtot_R = 0.0;
    for (int k = 1; k <= nk; ++k) {
        const double uu = col.uk(k);
        double f0, h0m, fL, hLm, fR, hRm;
        col.eval(k, uu + du0, f0, h0m);
        col.eval(k, uu + duL, fL, hLm);
        col.eval(k, uu + duR, fR, hRm);
        FAmt_0 = FAmt_0 + col.dy * h0m;
        FAmt_L = FAmt_L + col.dy * hLm;
        FAmt_R = FAmt_R + col.dy * hRm;
        uhtot_L = uhtot_L + fL;
        uhtot_R = uhtot_R + fR;
        // zonal_flux_thickness (:602-735) at u_cor = u + du, marginal faces
        double hm_cor = h0m;
        if (du0 != du) {
            double fc;
            col.eval(k, uu + du, fc, hm_cor);
        }
        bt_h_u[col.cu + (long long)(k - 1) * d.s2] = hm_cor;
    }
    double FA_0 = FAmt_0, FA_avg = FAmt_0;
    if ((duL - du0) != 0.0) FA_avg = uhtot_L / (duL - du0);
    if (FA_avg > dmax(FA_0, FAmt_L)) FA_avg = dmax(FA_0, FAmt_L);
    else if (FA_avg < dmin(FA_0, FAmt_L)) FA_0 = FA_avg;
    FA_u_W0[c2] = FA_0;
    FA_u_WW[c2] = FAmt_L;
    uBT_WW[c2] = (fabs(FA_0 - FAmt_L) <= (double)1e-12f * FA_0)
                     ? 0.0
                     : (1.5 * (duL - du0)) * ((FAmt_L - FA_avg) / (FAmt_L - FA_0));
    FA_0 = FAmt_0;
    FA_avg = FAmt_0;
    if ((duR - du0) != 0.0) FA_avg = uhtot_R / (duR - du0);
    if (FA_avg > dmax(FA_0, FAmt_R)) FA_avg = dmax(FA_0, FAmt_R);
    else if (FA_avg < dmin(FA_0, FAmt_R)) FA_0 = FA_avg;
    FA_u_E0[c2] = FA_0;
    FA_u_EE[c2] = FAmt_R;
    uBT_EE[c2] = (fabs(FAmt_R - FA_0) <= (double)1e-12f * FA_0)
                     ? 0.0
                     : (1.5 * (duR - du0)) * ((FAmt_R - FA_avg) / (FAmt_R - FA_0));
}

// ----------------------------------------------------- block column path ---
// A block of CW wavefronts owns CF = 64 consecutive faces I0-1 .. I0+62
// (one face per lane, the same face in every wavefront) and all their levels:
// wavefront w holds levels w*K+1 .. w*K+K of its lane's face in registers (u,
// the PPM edges and thickness of both neighbouring cells, the layer transport
// and marginal thickness of the last evaluation).  Every pass over the columns
// (first fluxes, each Newton evaluation, the BT_cont evaluations) evaluates all
// levels in parallel from registers, leaves the per-level terms in LDS, and
// then one wavefront per column sum adds them up in the reference's k order,
// one lane per face -- a sequential sum costs 64 faces one add per level.
// The Newton logic of a face runs redundantly in every wavefront on the same
// values, so all copies agree; the block keeps iterating while any face
// needs another evaluation.  The block updates the 63 cells I0 .. I0+62, whose
// convergence needs the transport at the face to the west (the neighbouring
// lane); neighbouring blocks overlap by one face, which both evaluate
// identically and only one writes.
constexpr int CF = 64;


__device__ __forceinline__ double vabs(double x) { return fabs(x); }
__device__ __forceinline__ float vabs(float x) { return fabsf(x); }
__device__ __forceinline__ double vsign1(double s) { return copysign(1.0, s); }
__device__ __forceinline__ float vsign1(float s) { return copysignf(1.0f, s); }
template <typename T>
__device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }

// max/min as single instructions.  The reference's MAX/MIN differ from them
// only for NaN operands and for a pair of zeros of opposite sign; the PPM
// slope below applies them to thicknesses (never -0: a layer thickness is
// positive or the +0 that continuity's max(h, 0) leaves) and to differences
// x - y with x >= y and absolute values (>= +0, never -0).
__device__ __forceinline__ double hmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double hmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float hmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float hmin(float a, float b) { return __builtin_fminf(a, b); }

// PPM_reconstruction_x (:1124-1247) slope from the three cells, branch-free
template <typename T>
__device__ __forceinline__ T ppm_slope_v(T mL, T m0, T mR, T hm, T h0, T hp)
{
    const T s = T(0.5) * (hp - hm);
    const T dMx = hmax(hmax(hp, hm), h0) - h0;
    const T dMn = h0 - hmin(hmin(hp, hm), h0);
    const T slp = vsign1(s) * hmin(vabs(s), T(2.0) * hmin(dMx, dMn));
    return ((mL * m0 * mR) == T(0.0)) ? T(0.0) : slp;
}

// Unlimited PPM edges of the cell at offset q of h6/m6 (cells i-2 .. i+3)
template <typename T>
__device__ __forceinline__ void ppm_edges_v(const T *m6, const T *h6, const T *sl, int q, T &hl, T &hr)
{
    const T hi = h6[q];
    const T mm = m6[q - 1], mp = m6[q + 1];
    const T h_im1 = mm * h6[q - 1] + (T(1.0) - mm) * hi;
    const T h_ip1 = mp * h6[q + 1] + (T(1.0) - mp) * hi;
    const T oneSixth = T(1.0) / T(6.0);
    hl = T(0.5) * (h_im1 + hi) + oneSixth * (sl[q - 2] - sl[q - 1]);
    hr = T(0.5) * (h_ip1 + hi) + oneSixth * (sl[q - 1] - sl[q]);
}

// PPM_limit_pos (:1253-1294) / PPM_limit_CW84 (:1298-1338) on one cell, with
// the reference's REAL(4) locals
__device__ __forceinline__ void ppm_limit(int monotonic, double hi, double h_min, double &hl, double &hr)
{
    if (monotonic) {
        const float h_i = (float)hi;
        if ((hr - (double)h_i) * ((double)h_i - hl) <= 0.0) {
            hl = (double)h_i;
            hr = (double)h_i;
        } else {
            const float RLdiff = (float)(hr - hl);
            const float RLmean = (float)(0.5 * (hr + hl));
            const float FunFac = 6.0f * RLdiff * (h_i - RLmean);
            const float RLdiff2 = RLdiff * RLdiff;
            if (FunFac > RLdiff2) hl = (double)(3.0f * h_i) - 2.0 * hr;
            if (FunFac < -RLdiff2) hr = (double)(3.0f * h_i) - 2.0 * hl;
        }
        return;
    }
    const float curv = (float)(3.0 * ((hl + hr) - 2.0 * hi));
    if ((double)curv > 0.0) {
        const float dh = (float)(hr - hl);
        if (fabsf(dh) < curv) {
            if (hi <= h_min) {
                hl = hi;
                hr = hi;
            } else {
                const double lhs = (double)(12.0f * curv) * (hi - h_min);
                const float rhs = curv * curv + 3.0f * (dh * dh);
                if (lhs < (double)rhs) {
                    const float scale = (float)(lhs / (double)rhs);
                    hl = hi + (double)scale * (hl - hi);
                    hr = hi + (double)scale * (hr - hi);
                }
            }
        }
    }
}

// The same limiters entirely in FP32 (the mixed-precision form)
__device__ __forceinline__ void ppm_limit(int monotonic, float hi, float h_min, float &hl, float &hr)
{
    if (monotonic) {
        if ((hr - hi) * (hi - hl) <= 0.0f) {
            hl = hi;
            hr = hi;
        } else {
            const float RLdiff = hr - hl;
            const float RLmean = 0.5f * (hr + hl);
            const float FunFac = 6.0f * RLdiff * (hi - RLmean);
            const float RLdiff2 = RLdiff * RLdiff;
            if (FunFac > RLdiff2) hl = 3.0f * hi - 2.0f * hr;
            if (FunFac < -RLdiff2) hr = 3.0f * hi - 2.0f * hl;
        }
        return;
    }
    const float curv = 3.0f * ((hl + hr) - 2.0f * hi);
    if (curv > 0.0f) {
        const float dh = hr - hl;
        if (fabsf(dh) < curv) {
            if (hi <= h_min) {
                hl = hi;
                hr = hi;
            } else {
                const float lhs = (12.0f * curv) * (hi - h_min);
                const float rhs = curv * curv + 3.0f * (dh * dh);
                if (lhs < rhs) {
                    const float scale = lhs / rhs;
                    hl = hi + scale * (hl - hi);
                    hr = hi + scale * (hr - hi);
                }
            }
        }
    }
}

// The per-level terms of zonal_flux_layer (:516-598) that do not depend on
// the velocity, formed once per level from the limited PPM edges of the two
// cells beside the face, for each upwind side: p (u > 0, cell 0 upwind) and
// n (u < 0, cell 1).  A = the upwind edge facing the face, D = far edge - A,
// C = curv_3 = (far edge + A) - 2 h: the reference's expressions, evaluated
// once instead of in every pass of the Newton iteration (same operations on
// the same operands, so the same values).
template <typename T, int K>
struct ColRegs {
    T Ap[K], Dp[K], Cp[K], An[K], Dn[K], Cn[K];
    T dt, dy, IdxT0, IdxT1;
    __device__ __forceinline__ void set(int kk, T hW0, T hE0, T h0, T hW1, T hE1, T h1)
    {
        Ap[kk] = hE0;
        Dp[kk] = hW0 - hE0;
        Cp[kk] = (hW0 + hE0) - T(2.0) * h0;
        An[kk] = hW1;
        Dn[kk] = hE1 - hW1;
        Cn[kk] = (hE1 + hW1) - T(2.0) * h1;
    }
    // the layer transport uh and marginal thickness hm at velocity u: the
    // upwind side selected instead of branched; CFL = |c| (dt > 0 and
    // IdxT > 0, so c has the sign of u and |c| is the reference's
    // (u > 0 ? c : -c)); u = 0 (or NaN) gives uh = 0 and hm = (hW1 + hE0)/2,
    // handled in a branch that a wave takes only if one of its lanes has it.
    __device__ __forceinline__ void eval(int kk, T u, T &uh, T &hm) const
    {
        const bool pos = u > T(0.0);
        const T A = pos ? Ap[kk] : An[kk];
        const T D = pos ? Dp[kk] : Dn[kk];
        const T curv_3 = pos ? Cp[kk] : Cn[kk];
        const T c = u * dt * (pos ? IdxT0 : IdxT1);
        const T CFL = vabs(c);
        uh = (dy * u) * (A + CFL * (T(0.5) * D + curv_3 * (CFL - T(1.5))));
        hm = A + CFL * (D + T(3.0) * curv_3 * (CFL - T(1.0)));
        const bool zero = !(u > T(0.0)) && !(u < T(0.0));
        if (__any(zero)) {
            if (zero) {
                uh = T(0.0);
                hm = T(0.5) * (An[kk] + Ap[kk]);
            }
        }
    }
    // Marks the stored column as rewritten, so that the compiler re-derives the
    // per-level terms of the flux in every pass instead of keeping them live
    // across passes (which would double the registers per level).
    __device__ inline void fence()
    {
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
            asm volatile("" : "+v"(Ap[kk]), "+v"(Dp[kk]), "+v"(Cp[kk]), "+v"(An[kk]), "+v"(Dn[kk]), "+v"(Cn[kk]));
    }
};

// LDS of one block: U = u of every level (rows past nk hold level nk or
// nothing and are never read as values; u lives here rather than in
// registers), S0/S1 = per-level terms of a pass,
// R = results of the column sums (one row per quantity; rows 5..8 keep the
// Newton iteration's starting point for a second iteration).
template <typename T>
constexpr int walk_batch = 4;   // rows per batch of a column walk (two batches in flight)
// FW = faces per block (64: one level group per wave; 32: two, lanes 0-31
// and 32-63 holding the same faces at different levels).
template <typename T, int CW, int K, int FW = CF>
struct ContShared {
    static constexpr int L = CW * (64 / FW) * K;
    T U[L][FW], S0[L][FW], S1[L][FW], R[9][FW];
    // per-face 2D values, loaded with the block's set-up loads and parked here
    // until they are used rather than held in registers: P[0] = dt*IareaT of
    // the lane's cell (x), P[1] = dxCu of its face, P[2] = its barotropic
    // transport target, P[3] = dt*min(IareaT) of its two cells, P[4], P[5] =
    // dxT of cells 0 and 1; P[6], P[7] = du_eval and sum_dyhm of the face's
    // Newton iteration, kept over the output stores for BT_cont
    T P[8][FW];
};

// Visits rows 0..nk-1 (nk <= L) of one LDS column of ContShared in k order: step(x_k).
// Whole batches of B rows first, each batch's reads issued before the batch
// ahead of it is folded (two batches in flight, so the LDS latency overlaps
// the dependent chain; the two buffers alternate, so no copies wait on the
// reads); then the last nk % B (< B) rows one by one.  A remainder entered
// part-way through an unrolled sequence (a switch that falls through) was
// miscompiled for one entry point by the compiler of this toolchain: a
// hoisted operand of the step was left undefined on that path.
template <int L, typename T, int FW, typename Step>
__device__ __forceinline__ void lds_walk(const T (*col)[FW], int nk, int lane, Step step)
{
    constexpr int B = walk_batch<T>;
    int k0 = 0;
    if (nk >= B) {
        T x[B], y[B];
        auto load = [&](T(&v)[B], int r) __attribute__((always_inline)) {
#pragma unroll
            for (int b = 0; b < B; ++b) v[b] = col[r + b][lane];
        };
        auto fold = [&](const T(&v)[B]) __attribute__((always_inline)) {
#pragma unroll
            for (int b = 0; b < B; ++b) step(v[b]);
        };
        load(x, 0);   // x: rows k0 .. k0+B-1
        for (; k0 + 3 * B <= nk; k0 += 2 * B) {
            load(y, k0 + B);
            fold(x);
            load(x, k0 + 2 * B);
            fold(y);
        }
        if (k0 + 2 * B <= nk) {
            load(y, k0 + B);
            fold(x);
            fold(y);
            k0 += 2 * B;
        } else {
            fold(x);
            k0 += B;
        }
    }
#pragma nounroll
    for (; k0 < nk; ++k0) step(col[k0][lane]);
}

// Barrier of the block column path.  The waves of a block exchange data
// through LDS only (nothing a block stores to global memory is read back by
// it), so the barrier orders LDS operations alone: __syncthreads' workgroup
// fence would also wait for every global store still in flight, a memory
// round trip at the first barrier after the output stores.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Block-wide "any face has pred", for the predicates of the Newton iteration
// and of BT_cont: every wave holds all the block's faces (FW = 32: twice) and
// evaluates pred for each from the same operands (the column sums and parked
// 2D values in LDS, and what every wave derives from them by the same
// instructions), so one wave's vote is the block's and needs no barrier (an
// LDS vote and a barrier per test measured 4% of the kernel).  The caller's
// LDS reads and writes around a test are ordered by the barriers of the
// column sums.
__device__ __forceinline__ bool block_any(bool pred)
{
    return __any(pred);
}

// Sequential fold acc = op(acc, x_k) over rows 0..nk-1 of one LDS column
template <int L, typename T, int FW, typename Op>
__device__ __forceinline__ T lds_fold(const T (*col)[FW], int nk, int lane, T acc, Op op)
{
    lds_walk<L>(col, nk, lane, [&](T x) __attribute__((always_inline)) { acc = op(acc, x); });
    return acc;
}

// Column sum x0 + sum_k col[k] in k order: plain FP64 adds (the reference's
// summation), or for FP32 terms a Kahan-compensated sum.
template <int L, int FW>
__device__ __forceinline__ double col_sum(const double (*col)[FW], int nk, int lane, double x0)
{
    return lds_fold<L>(col, nk, lane, x0, [](double a_, double x_) { return a_ + x_; });
}

// the same with each term scaled by w first: x0 + sum_k (w * col[k])
template <int L, int FW>
__device__ __forceinline__ double col_sum(const double (*col)[FW], int nk, int lane, double x0, double w)
{
    return lds_fold<L>(col, nk, lane, x0, [w](double a_, double x_) { return a_ + w * x_; });
}

template <int L, int FW>
__device__ __forceinline__ float col_sum(const float (*col)[FW], int nk, int lane, float x0, float w = 1.0f,
                                         bool scaled = false)
{
    float acc = x0, c = 0.0f;
    lds_walk<L>(col, nk, lane, [&](float x) __attribute__((always_inline)) {
        const float y = (scaled ? w * x : x) - c;
        const float t = acc + y;
        c = (t - acc) - y;
        acc = t;
    });
    return acc;
}

// Y = false: the zonal pass.  A block covers the faces I0-1 .. I0+62 of row j
// and also updates the thickness of cells I0 .. I0+62 (the convergence needs
// the neighbouring lane's transport).  Y = true: the meridional pass, the same
// algorithm across the rows: a block covers the faces (i, J), i = i0 .. i0+63,
// of one face row J (lanes still run along i, so every load stays coalesced);
// its convergence is a separate kernel (the faces J-1 belong to another block).
//
// FW = 32: a block covers 32 faces and each wave two level groups (lanes
// 0-31 and 32-63), so a block needs half the registers and LDS of FW = 64 and
// two blocks share a CU: one block's loads, folds and barriers overlap the
// other's flux passes.  Both halves of a wave run the Newton logic of the same
// face (the same inputs, so the same decisions); the column folds are done by
// whole waves, both halves folding the same column (no divergence) and
// storing the same value.
//
// FP32 form: the Newton stopping test of zonal_flux_adjust also accepts a
// residual at the rounding level of the layer fluxes (4 ulp of sum |uh_k| of
// the first pass); the FP64 tolerance of 1e-12 m lies below what FP32 sums
// resolve, and the iteration would otherwise run to its 20-pass limit.
template <typename T, bool Y, int CW, int K, int FW, bool EVEN>
__device__ __forceinline__ void continuity_body(const ContArgsT<T> &a, int monotonic, unsigned bx, unsigned by,
                                                bool pf_on = false)
{
    extern __shared__ double smem_raw[];
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int L = ContShared<T, CW, K, FW>::L;
    ContShared<T, CW, K, FW> &S = *reinterpret_cast<ContShared<T, CW, K, FW> *>(smem_raw);
    const Dims &d = a.d;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & (FW - 1);                  // the face of the block this thread works on
    const int grp = w * (64 / FW) + (int)((threadIdx.x & 63) / FW);   // its level group
    // face (fi, fj) between cell 0 = (fi, fj) and cell 1 = (fi+1, fj) [x] or (fi, fj+1) [y]
    // x: a block advances by FW-2 faces (its last face lane only lends its cell
    // to the lane below; its first lane's face feeds the convergence of the second)
    const int f_raw = Y ? d.isc + bx * FW + lane : d.isc - 1 + bx * (FW - 2) + lane;
    const int f_max = d.iec;
    const int fi = f_raw > f_max ? f_max : f_raw;           // clamped: evaluated, never written
    const int fj = Y ? d.jsc - 1 + (int)by : d.jsc + (int)by;
    const bool live = f_raw <= f_max;
    auto cell = [&](int q) __attribute__((always_inline)) -> long long {   // q-th cell of the stencil, q = 2: cell 0
        return Y ? IX2(d, fi, fj - 2 + q) : IX2(d, fi - 2 + q, fj);
    };
    const int i = fi, j = fj;
    const int nk = d.nk;
    const long long s = d.s2;
    // my levels.  EVEN = false: k0+1 .. k0+K, every slot evaluated (a slot
    // past nk evaluates level nk into an LDS row past nk, which no sum or
    // output reads).  EVEN = true
    // (one level group per wave and CW*(K-1) <= nk <= CW*K, which the
    // launcher checks): the nk levels spread evenly over the waves, K of them
    // for the first nk - CW*(K-1) waves and K-1 for the rest, so that the
    // SIMDs evaluate about nk/4 levels each per pass instead of CW*K/4 (at nk
    // = 75 with 12 waves of 7: 19 instead of 21); only a wave's last slot can
    // be empty, skipped in a branch on one wave-uniform flag.
    int k0;
    bool tail = true;
    if constexpr (EVEN) {
        static_assert(FW == 64, "one level group per wave");
        const int g = __builtin_amdgcn_readfirstlane(grp);
        const int full = nk - CW * (K - 1);   // waves with K levels
        tail = g < full;
        k0 = g * (K - 1) + (tail ? g : full);
    } else {
        k0 = grp * K;
    }
    auto valid = [&](int kk) __attribute__((always_inline)) {
        if constexpr (EVEN) return kk < K - 1 || tail;
        else return k0 + kk < nk;
    };
    // slots whose terms are formed: EVEN skips its empty slot
    auto formed = [&](int kk) __attribute__((always_inline)) { return !EVEN || valid(kk); };

    ColRegs<T, K> col;
    col.dt = (T)a.dt;
    const long long c2 = IX2(d, fi, fj);
    col.dy = (T)a.dy_Cu[c2];
    if constexpr (!Y && FW == 64) {
        // cell 1's IdxT is the next lane's cell 0's, read at cell
        // min(f_raw, iec+1): the lane after the last live one holds cell
        // iec+1 (a lane past iec evaluates values that are never written)
        const int ci0 = f_raw < d.iec + 1 ? f_raw : d.iec + 1;
        col.IdxT0 = (T)a.IdxT[IX2(d, ci0, fj)];
        col.IdxT1 = __shfl_down(col.IdxT0, 1, FW);
    } else {
        col.IdxT0 = (T)a.IdxT[cell(2)];
        col.IdxT1 = (T)a.IdxT[cell(3)];
    }
    T mCu = T(0.0);   // read by the CFL folds of waves 2 and 3 only
    if (w == 2 || w == 3) mCu = (T)a.mask2dCu[c2];
    const long long cu = IX3(d, fi, fj, 1);
    // the 2D values read after the column passes, loaded by wave 0 alone
    // (which parks them in LDS for every wave) and issued with the column
    // loads: a load issued where its value is first used costs the whole
    // block a memory round trip at that point
    double IareaT0 = 0.0, IareaT1 = 0.0, dxT0 = 0.0, dxT1 = 0.0, uhbt_f = 0.0, dxCu_f = 0.0;
    if (w == 0) {
        IareaT0 = a.IareaT[cell(2)];
        IareaT1 = a.IareaT[cell(3)];
        dxT0 = a.dxT[cell(2)];
        dxT1 = a.dxT[cell(3)];
        uhbt_f = a.uhbt ? a.uhbt[c2] : 0.0;
        dxCu_f = a.dxCu[c2];
    }
    auto park = [&]() __attribute__((always_inline)) {
        if (w == 0) {
            if (!Y) S.P[0][lane] = (T)(a.dt * IareaT0);
            S.P[1][lane] = (T)dxCu_f;
            S.P[2][lane] = (T)uhbt_f;
            S.P[3][lane] = (T)(a.dt * dmin(IareaT0, IareaT1));
            S.P[4][lane] = (T)dxT0;
            S.P[5][lane] = (T)dxT1;
        }
    };
    // PPM edges of cells 0 and 1 (zonal_edge_thickness :185-207); the loads
    // of a slot past the lane's levels read level min(k, nk) (in bounds) and
    // nothing is formed from them.  x: each lane reconstructs its own cell 0 from the
    // 5-cell stencil and takes cell 1 from the next lane (whose cell 0 it is),
    // so every cell is reconstructed once.  y: both cells from the 6-cell
    // stencil along j.
    if (!Y) {
        const int ci = f_raw < d.iec + 1 ? f_raw : d.iec + 1;   // this lane's cell (in bounds)
        auto ccell = [&](int q) __attribute__((always_inline)) -> long long { return IX2(d, ci - 2 + q, fj); };
        // STAGE: each level's row of thicknesses is loaded once, the lanes'
        // own cells (position t = lane holds cell f_raw(lane 0) - 2 + t) and
        // the four beyond the last lane (t = 64..67, those of all the wave's
        // levels in one load), and each lane reads its 5-cell stencil t =
        // lane .. lane+4 back from LDS: one load per level and lane instead
        // of five (a load with few active lanes costs about what a full one
        // does, so the extensions share one).  The rows are
        // staged in the wave's own rows of S0 (levels 0..KA-1) and S1 (the
        // rest), which no other wave touches before the first pass's barrier
        // and this wave's first pass overwrites only after these reads.  A
        // cell index past iec+3 (the last one the lanes of the unstaged form
        // read) is clamped there: only lanes past iec+1, which are evaluated
        // and never written, read a clamped cell.  STAGE_M: the row of masks
        // likewise, after the thickness rows in S1 where they fit.
        constexpr bool STAGE = FW == 64 && K >= 3;
        constexpr int KA = K / 2, RL = FW + 4;
        static_assert(!STAGE || (KA * RL <= (K - 1) * FW && (K - KA) * RL <= K * FW &&
                                 (K - 1 - KA) * RL <= (K - 1) * FW),
                      "staged rows exceed the wave's own rows");
        constexpr bool STAGE_M = STAGE && (K - KA + 1) * RL <= (K - 1) * FW;
        T m5[5], hs[K][5], us[K];
        if constexpr (!STAGE_M) {
#pragma unroll
            for (int q = 0; q < 5; ++q) m5[q] = (T)a.mask2dT[ccell(q)];
        }
        if constexpr (STAGE) {
            const int c_lo = f_raw - 2, c_hi = d.iec + 3;
            const long long rA = IX2(d, c_lo < c_hi ? c_lo : c_hi, fj);
            const long long rB = IX2(d, c_lo + FW < c_hi ? c_lo + FW : c_hi, fj);
            T mA = T(0.0), mB = T(0.0);
            if constexpr (STAGE_M) {
                mA = (T)a.mask2dT[rA];
                if (lane < 4) mB = (T)a.mask2dT[rB];
            }
            T hA[K], hX = T(0.0);
            // the four cells past the last lane of every level in one load:
            // lane 4*kk + x holds cell 64 + x of slot kk
            const int xk = lane >> 2;
            const int cx = f_raw - lane - 2 + FW + (lane & 3);
            const long long rX = IX2(d, cx < c_hi ? cx : c_hi, fj);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                const int kr = k0 + kk + 1;
                const int k = kr < nk ? kr : nk;
                const long long o = (long long)(k - 1) * s;
                us[kk] = a.u[cu + o];
                if (!formed(kk)) continue;
                hA[kk] = a.hin[rA + o];
            }
            if (lane < 4 * K) {
                const int kr = k0 + xk + 1;
                hX = a.hin[rX + (long long)((kr < nk ? kr : nk) - 1) * s];
            }
            park();
            auto row = [&](int kk) __attribute__((always_inline)) -> T * {
                return kk < KA ? &S.S0[k0][0] + kk * RL : &S.S1[k0][0] + (kk - KA) * RL;
            };
            T *const mrow = &S.S1[k0][0] + (K - KA) * RL;
            if constexpr (STAGE_M) {
                mrow[lane] = mA;
                if (lane < 4) mrow[FW + lane] = mB;
            }
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                if (!formed(kk)) continue;
                row(kk)[lane] = hA[kk];
            }
            if (lane < 4 * K && formed(xk)) {
                T *const r = xk < KA ? &S.S0[k0][0] + xk * RL : &S.S1[k0][0] + (xk - KA) * RL;
                r[FW + (lane & 3)] = hX;
            }
            // one wave's LDS operations complete in order: the fences only
            // keep the compiler from moving the reads above the stores
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            if constexpr (STAGE_M) {
#pragma unroll
                for (int q = 0; q < 5; ++q) m5[q] = mrow[lane + q];
            }
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                if (!formed(kk)) continue;
#pragma unroll
                for (int q = 0; q < 5; ++q) hs[kk][q] = row(kk)[lane + q];
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                const int kr = k0 + kk + 1;
                const int k = kr < nk ? kr : nk;
                const long long o = (long long)(k - 1) * s;
                us[kk] = a.u[cu + o];
#pragma unroll
                for (int q = 0; q < 5; ++q) hs[kk][q] = a.hin[ccell(q) + o];
            }
            park();
        }
        T wW[K], wE[K], w0[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (!formed(kk)) continue;
            S.U[k0 + kk][lane] = us[kk];
            const T *h5 = hs[kk];
            T sl[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) sl[q] = ppm_slope_v(m5[q], m5[q + 1], m5[q + 2], h5[q], h5[q + 1], h5[q + 2]);
            ppm_edges_v(m5, h5, sl, 2, wW[kk], wE[kk]);
            w0[kk] = h5[2];
        }
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (!formed(kk)) continue;
            ppm_limit(monotonic, w0[kk], (T)a.h_min, wW[kk], wE[kk]);
            col.set(kk, wW[kk], wE[kk], w0[kk], __shfl_down(wW[kk], 1, FW), __shfl_down(wE[kk], 1, FW),
                    __shfl_down(w0[kk], 1, FW));
        }
    } else {
        T m6[6], hs[K][6], us[K];
#pragma unroll
        for (int q = 0; q < 6; ++q) m6[q] = (T)a.mask2dT[cell(q)];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const int kr = k0 + kk + 1;
            const int k = kr < nk ? kr : nk;
            const long long o = (long long)(k - 1) * s;
            us[kk] = a.u[cu + o];
#pragma unroll
            for (int q = 0; q < 6; ++q) hs[kk][q] = a.hin[cell(q) + o];
        }
        park();
        T wW[K], wE[K], w0[K], eW[K], eE[K], e0[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (!formed(kk)) continue;
            S.U[k0 + kk][lane] = us[kk];
            const T *h6 = hs[kk];
            T sl[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) sl[q] = ppm_slope_v(m6[q], m6[q + 1], m6[q + 2], h6[q], h6[q + 1], h6[q + 2]);
            ppm_edges_v(m6, h6, sl, 2, wW[kk], wE[kk]);
            ppm_edges_v(m6, h6, sl, 3, eW[kk], eE[kk]);
            w0[kk] = h6[2];
            e0[kk] = h6[3];
        }
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (!formed(kk)) continue;
            ppm_limit(monotonic, w0[kk], (T)a.h_min, wW[kk], wE[kk]);
            ppm_limit(monotonic, e0[kk], (T)a.h_min, eW[kk], eE[kk]);
            col.set(kk, wW[kk], wE[kk], w0[kk], eW[kk], eE[kk], e0[kk]);
        }
    }

    auto uk = [&](int kk) __attribute__((always_inline)) -> T { return S.U[k0 + kk][lane]; };
    // one pass: evaluate every level at u + du; terms (layer transport, hm) to
    // LDS.  S0 keeps each lane's transports of its last evaluation, the
    // result the face outputs and the convergence read.
    // Each level's u is read before the stores of the level above it, so the
    // LDS read overlaps that level's arithmetic (EVEN: the read for a wave's
    // empty slot stays inside U and its value is not used).
    auto eval_pass = [&](T du, bool on) __attribute__((always_inline)) {
        if (!on) return;
        col.fence();
        T u_next = uk(0);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const T u_k = u_next;
            if (kk + 1 < K) u_next = uk(kk + 1);
            if (!formed(kk)) continue;
            T f, hm;
            col.eval(kk, u_k + du, f, hm);
            S.S0[k0 + kk][lane] = f;
            S.S1[k0 + kk][lane] = hm;   // scaled by dy in the column sum
        }
    };
    // x1 + sum_k dy*hm(k): S1 holds the marginal thickness hm of the pass, so
    // that the last Newton pass also yields BT_cont's h_u without a pass of its own
    auto sum_dy = [&](T x1) __attribute__((always_inline)) -> T {
        if constexpr (sizeof(T) == 8) return col_sum<L>(S.S1, nk, lane, x1, col.dy);
        else return col_sum<L>(S.S1, nk, lane, x1, col.dy, true);
    };
    // column sums in k order by wavefronts 0 and 1 (R[0] = x0 + sum S0, R[1] = x1 + sum dy*S1)
    auto sum_pass = [&](T x0, T x1, bool on) __attribute__((always_inline)) {
        lds_barrier();
        if (w < 2 && on)
            S.R[w][lane] = (w == 0) ? col_sum<L>(S.S0, nk, lane, x0) : sum_dy(x1);
        lds_barrier();
    };

    // first layer fluxes, their sums and the CFL bracket (zonal_mass_flux :300-393)
    const T CFL_dt = (T)a.CFL_limit_adjust / (T)a.dt;
    eval_pass(T(0.0), true);
    lds_barrier();
    // Prefetch (pf_on: the single-model launch, whose blocks are dispatched
    // in blockIdx order, round-robin over the 8 XCDs): the u and h rows of
    // block b + 256, which runs on this block's XCD (256 is a multiple of 8)
    // about one block lifetime later at one block per CU.  One 4-byte load per
    // lane touches one 128-byte line of a row (9 lines per level in FP64, 5 in
    // FP32, as many levels as the block's lanes cover).  A block's set-up
    // waits on its own loads (the last wave of each SIMD receives its rows
    // about 14 K cycles after the block starts); lines this prefetch brought
    // in arrive sooner (0.98 against 1.017 ms per FP64 launch at
    // 720x720x75; the FP32 launch gains 4%, and a distance of 512 measured no
    // better for it).  The loads are issued after this block's first pass,
    // their value is consumed by an empty asm at the end so that they are
    // kept, and they read inside the arrays, at indices that the block they
    // prefetch for reads.
    int pf = 0;
    if constexpr (!Y && FW == 64) {
        const unsigned gx = gridDim.x, b2 = bx + gx * by + 256;
        if (pf_on && b2 < gx * gridDim.y) {
            constexpr int LU = (64 * (int)sizeof(T) + 127) / 128, LH = (68 * (int)sizeof(T) + 127) / 128;
            constexpr int ST = 128 / (int)sizeof(T);
            const int bx2 = (int)(b2 % gx), by2 = (int)(b2 / gx);
            const int q = w * 64 + (int)(threadIdx.x & 63);
            const int lvl = q / (LU + LH), part = q % (LU + LH);
            if (lvl < nk) {
                const int f0 = d.isc - 1 + bx2 * (FW - 2), fj2 = d.jsc + by2;
                const int *pp;
                if (part < LU) {
                    const int f = f0 + part * ST;
                    pp = (const int *)(a.u + IX3(d, f < d.iec ? f : d.iec, fj2, lvl + 1));
                } else {
                    const int c = f0 - 2 + (part - LU) * ST;
                    pp = (const int *)(a.hin + IX2(d, c < d.iec + 3 ? c : d.iec + 3, fj2) + (long long)lvl * s);
                }
                pf = *pp;
            }
        }
    }
    const T dx_W = S.P[4][lane], dx_E = S.P[5][lane];
    if (w < 2) {
        S.R[w][lane] = (w == 0) ? col_sum<L>(S.S0, nk, lane, T(0.0)) : sum_dy(T(0.0));
    } else if (w == 2) {
        S.R[2][lane] = lds_fold<L>(S.U, nk, lane, T(2.0) * (CFL_dt * dx_W) * T(1.0), [&](T x, T uu) {
            return (x > dx_W * CFL_dt - uu * mCu) ? (dx_W * CFL_dt - uu) : x;
        });
    } else if (w == 3) {
        S.R[3][lane] = lds_fold<L>(S.U, nk, lane, -(T(2.0) * (CFL_dt * dx_E) * T(1.0)), [&](T x, T uu) {
            return (x < -dx_E * CFL_dt - uu * mCu) ? -(dx_E * CFL_dt + uu) : x;
        });
    } else if (F32 && w == 4) {   // scale of the rounding of the flux sum
        S.R[4][lane] = lds_fold<L>(S.S0, nk, lane, T(0.0), [](T x, T f) { return x + vabs(f); });
    }
    lds_barrier();
    // the Newton iteration's starting point: the first pass's sums and the CFL
    // bracket; LDS rows 5..8 keep it for the zero-target iteration of BT_cont
    // (so that no register holds it across the first iteration)
    struct Start {
        T uh_tot_0, duhdu_tot_0, du_max_CFL, du_min_CFL;
    };
    const Start start{S.R[0][lane], S.R[1][lane], vmax(S.R[2][lane], T(0.0)), vmin(S.R[3][lane], T(0.0))};
    if (w == 0) {
        S.R[5][lane] = start.uh_tot_0;
        S.R[6][lane] = start.duhdu_tot_0;
        S.R[7][lane] = start.du_max_CFL;
        S.R[8][lane] = start.du_min_CFL;
    }
    const T bt = S.P[2][lane];
    const T tol_floor = F32 ? S.P[3][lane] * (T(4.0f * 1.1920929e-7f) * S.R[4][lane]) : T(0.0);
    const T tol_eta0 = (T)a.tol_eta0, tol_vel = (T)a.tol_vel;

    // zonal_flux_adjust (:739-916).  keep: uh_3d present (fluxes of the last
    // evaluation are the result; evaluated also at itt = max_itts).  Returns
    // du; *du_eval = du of the last evaluation that was summed (NaN if none)
    // and *dyhm_eval = sum_k dy*hm at u + *du_eval (the last duhdu_tot).
    auto flux_adjust = [&](const Start &st, T uhbt, bool keep, T *du_eval, T *dyhm_eval)
                           __attribute__((always_inline)) -> T {
        const int max_itts = 20;
        const T dtIarea = S.P[3][lane];
        T du = T(0.0), du_max = st.du_max_CFL, du_min = st.du_min_CFL;
        T uh_err = st.uh_tot_0 - uhbt, duhdu_tot = st.duhdu_tot_0;
        T uh_err_best = vabs(uh_err);
        // x: lanes past the last face and the block's last lane (no cell 1 of
        // its own) take no Newton passes
        bool do_I = Y || (live && lane < FW - 1);
        *du_eval = T(0.0);   // the first pass evaluated at du = 0
        for (int itt = 1; itt <= max_itts; ++itt) {
            if (do_I) {
                T tol_eta;
                if (itt <= 1) tol_eta = (T)1e-6f * tol_eta0;
                else if (itt == 2) tol_eta = (T)1e-4f * tol_eta0;
                else if (itt == 3) tol_eta = (T)1e-2f * tol_eta0;
                else tol_eta = tol_eta0;
                if (F32) tol_eta = vmax(tol_eta, tol_floor);
                if (uh_err > T(0.0)) du_max = du;
                else if (uh_err < T(0.0)) du_min = du;
                else do_I = false;
                if (do_I) {
                    const T aerr = vabs(uh_err);
                    if ((dtIarea * aerr > tol_eta) || ((aerr > tol_vel * duhdu_tot) || (aerr > uh_err_best))) {
                        const T ddu = -uh_err / duhdu_tot;
                        const T du_prev = du;
                        du = du + ddu;
                        if (vabs(ddu) < (T)1.0e-15f * vabs(du)) {
                            do_I = false;
                        } else if (ddu > T(0.0)) {
                            if (du >= du_max) {
                                du = T(0.5) * (du_prev + du_max);
                                if (du_max - du_prev < (T)1.0e-15f * vabs(du)) do_I = false;
                            }
                        } else {
                            if (du <= du_min) {
                                du = T(0.5) * (du_prev + du_min);
                                if (du_prev - du_min < (T)1.0e-15f * vabs(du)) do_I = false;
                            }
                        }
                    } else {
                        do_I = false;
                    }
                }
            }
            // a finished column keeps its du for the rest of the reference's row loop
            const bool ev = do_I && ((itt < max_itts) || keep);
            if (!block_any(ev)) break;
            eval_pass(du, ev);
            const bool summed = ev && itt < max_itts;
            sum_pass(-uhbt, T(0.0), summed);
            if (summed) {
                uh_err = S.R[0][lane];
                duhdu_tot = S.R[1][lane];
                uh_err_best = vmin(uh_err_best, vabs(uh_err));
                *du_eval = du;
            } else if (ev) {
                *du_eval = (T)NAN;
            }
        }
        *dyhm_eval = duhdu_tot;
        return du;
    };
    T du_eval, sum_dyhm;
    const T du = flux_adjust(start, bt, true, &du_eval, &sum_dyhm);
    // here S0 holds the layer fluxes of the last evaluation and, when
    // du_eval == du, sum_dyhm the sum of dy*hm at u + du
    const bool last_at_du = (du_eval == du);
    // every wave stores its own copy (all hold the same values), so each
    // reads back what it stored whether or not a barrier lies between
    S.P[6][lane] = du_eval;
    S.P[7][lane] = sum_dyhm;
    // the convergence's cell thicknesses (cell I = I0 + lane, every level of
    // the lane's group; clamped levels repeat nk), issued before the face
    // stores so that one round trip covers all levels
    T hc[K];
    if (!Y) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const int kr = k0 + kk + 1;
            hc[kk] = a.hin[IX3(d, i, j, kr < nk ? kr : nk)];
        }
    }

    // outputs at the face: uh, u_cor, du_cor; a face shared with the western
    // neighbour block is written by that block.  u_cor = u + du needs only u
    // (in LDS to the end) and du: with the BT_cont outputs its stores wait for
    // the CFL folds of BT_cont, where the other waves are idle, instead of
    // joining the burst of the other outputs (1.8% of the kernel)
    const bool writes_face = live && (Y || ((lane > 0 || bx == 0) && lane < FW - 1));
    auto store_u_cor = [&]() __attribute__((always_inline)) {
        if (a.u_cor && writes_face) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                if (valid(kk)) a.u_cor[cu + (long long)(k0 + kk) * s] = uk(kk) + du;
        }
    };
    if (writes_face) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (valid(kk)) a.uh[cu + (long long)(k0 + kk) * s] = S.S0[k0 + kk][lane];
        }
        if (a.du_cor && grp == 0) a.du_cor[c2] = (double)du;
    }
    if (!a.bt_h_u) store_u_cor();
    // continuity_zonal_convergence (:148-176) for cell i = I (faces I-1 and I)
    if (!Y) {
        const T dtIareaT = S.P[0][lane];   // dt*IareaT(i, j): cell 0 of the face
        const bool writes_cell = live && lane > 0 && lane < FW - 1 && i >= d.isc;
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (writes_cell && valid(kk)) {
                const long long c3 = IX3(d, i, j, k0 + kk + 1);
                const T fw = S.S0[k0 + kk][lane - 1];   // face I-1 (the lane to the west), same level
                const T v = hc[kk] - dtIareaT * (S.S0[k0 + kk][lane] - fw);
                a.h[c3] = vmax(v, T(0.0));
            }
        }
    }
    if (!a.bt_h_u) {   // block-uniform
        asm volatile("" ::"v"(pf));
        return;
    }

    // h_u = marginal thickness at u_cor = u + du (zonal_flux_thickness :602-735):
    // the last Newton pass's (in S1, each wave's own rows) when that was at
    // u + du, else evaluated here.  S1 holds until the next pass, so unless a
    // pass comes before the duL pass (the zero-target iteration, the du0
    // pass) the stores wait for the CFL folds, like u_cor's
    auto store_h_u = [&]() __attribute__((always_inline)) {
        if (!writes_face) return;
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (!valid(kk)) continue;
            T hm = S.S1[k0 + kk][lane];
            if (!last_at_du) {
                T fc;
                col.eval(kk, uk(kk) + du, fc, hm);
            }
            a.bt_h_u[cu + (long long)(k0 + kk) * s] = hm;
        }
    };
    bool h_u_stored = false;   // block-uniform

    // set_zonal_BT_cont (:920-1121).  The zero-transport correction equals du
    // whenever the target is zero (identical Newton iterations).
    const bool bt_any = a.uhbt && block_any(S.P[2][lane] != T(0.0));
    T du0 = du, du0_eval = S.P[6][lane], FAmt_0 = S.P[7][lane];
    if (bt_any) {   // without uhbt, bt = 0 everywhere
        store_h_u();
        h_u_stored = true;
        const T bt = S.P[2][lane];
        T de, s0;
        const Start again{S.R[5][lane], S.R[6][lane], S.R[7][lane], S.R[8][lane]};
        const T d2 = flux_adjust(again, T(0.0), false, &de, &s0);
        if (bt != T(0.0)) { du0 = d2; du0_eval = de; FAmt_0 = s0; }
    }
    // FAmt_0 = sum dy*hm(u + du0): the sum of the iteration's last evaluation
    // when that was at du0 (the same terms in the same order), else a pass of its own
    const bool have0 = (du0_eval == du0);
    if (block_any(!have0)) {
        if (!h_u_stored) store_h_u();
        h_u_stored = true;
        eval_pass(du0, !have0);
        sum_pass(T(0.0), T(0.0), !have0);
        if (!have0) FAmt_0 = S.R[1][lane];
    }
    // test velocities duR, duL (visc_rem_lim = 1)
    const T Idt = T(1.0) / (T)a.dt;
    const T CFL_min = (T)1e-6f;
    const T du_CFL = (CFL_min * Idt) * S.P[1][lane];
    lds_barrier();
    if (w >= 2) {   // while waves 0 and 1 fold
        store_u_cor();
        if (!h_u_stored) store_h_u();
    }
    if (w == 0) {
        S.R[2][lane] = lds_fold<L>(S.U, nk, lane, vmin(T(0.0), du0 - du_CFL), [&](T x, T uu) {
            return (uu + x > -du_CFL) ? -(uu + du_CFL) : x;
        });
    } else if (w == 1) {
        S.R[3][lane] = lds_fold<L>(S.U, nk, lane, vmax(T(0.0), du0 + du_CFL), [&](T x, T uu) {
            return (uu + x < du_CFL) ? -(uu - du_CFL) : x;
        });
    }
    lds_barrier();
    if (w < 2) {
        store_u_cor();
        if (!h_u_stored) store_h_u();
    }
    const T duR = S.R[2][lane], duL = S.R[3][lane];
    // FAmt_L, uhtot_L at u + duL; FAmt_R, uhtot_R at u + duR
    eval_pass(duL, true);
    sum_pass(T(0.0), T(0.0), true);
    const T uhtot_L = S.R[0][lane], FAmt_L = S.R[1][lane];
    eval_pass(duR, true);
    sum_pass(T(0.0), T(0.0), true);
    const T uhtot_R = S.R[0][lane], FAmt_R = S.R[1][lane];
    asm volatile("" ::"v"(pf));
    if (!writes_face || grp != 0) return;
    T FA_0 = FAmt_0, FA_avg = FAmt_0;
    if ((duL - du0) != T(0.0)) FA_avg = uhtot_L / (duL - du0);
    if (FA_avg > vmax(FA_0, FAmt_L)) FA_avg = vmax(FA_0, FAmt_L);
    else if (FA_avg < vmin(FA_0, FAmt_L)) FA_0 = FA_avg;
    a.FA_u_W0[c2] = (double)FA_0;
    a.FA_u_WW[c2] = (double)FAmt_L;
    a.uBT_WW[c2] = (double)((vabs(FA_0 - FAmt_L) <= (T)1e-12f * FA_0)
                                ? T(0.0)
                                : (T(1.5) * (duL - du0)) * ((FAmt_L - FA_avg) / (FAmt_L - FA_0)));
    FA_0 = FAmt_0;
    FA_avg = FAmt_0;
    if ((duR - du0) != T(0.0)) FA_avg = uhtot_R / (duR - du0);
    if (FA_avg > vmax(FA_0, FAmt_R)) FA_avg = vmax(FA_0, FAmt_R);
    else if (FA_avg < vmin(FA_0, FAmt_R)) FA_0 = FA_avg;
    a.FA_u_E0[c2] = (double)FA_0;
    a.FA_u_EE[c2] = (double)FAmt_R;
    a.uBT_EE[c2] = (double)((vabs(FAmt_R - FA_0) <= (T)1e-12f * FA_0)
                                ? T(0.0)
                                : (T(1.5) * (duR - du0)) * ((FAmt_R - FA_avg) / (FAmt_R - FA_0)));
}

template <typename T, bool Y, int CW, int K, int FW = CF, bool EVEN = false>
__global__ void __launch_bounds__(64 * CW) __attribute__((amdgpu_waves_per_eu(
    (sizeof(T) == 4 || FW == 32) ? 4 : (CW + 3) / 4, (CW >= 12 || sizeof(T) == 4 || FW == 32) ? 4 : 2)))
k_continuity_blk(ContArgsT<T> a, int monotonic)
{
    continuity_body<T, Y, CW, K, FW, EVEN>(a, monotonic, blockIdx.x, blockIdx.y, true);
}

// the members of an ensemble that one batched launch covers: entry q of its
// BatchMap is member v[q]
struct MemberIndex {
    int v[kMaxMembers];
};

// the members of an ensemble whose depth takes the block shape CW x K, in one
// launch (EVEN: every one of them spreads its levels evenly)
template <typename T, bool Y, int CW, int K, bool EVEN>
__global__ void __launch_bounds__(64 * CW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1, (CW >= 16 || sizeof(T) == 4) ? 4 : 2)))
k_continuity_batch(const ContArgsT<T> *__restrict__ args, BatchMap mp, MemberIndex mi)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const ContArgsT<T> a = args[mi.v[m]];
    continuity_body<T, Y, CW, K, CF, EVEN>(a, a.monotonic, local % mp.nbx[m], local / mp.nbx[m]);
}

// the driver's `h = h0` reset for the cells the convergence does not write:
// rows outside jsc..jec in full, and the isd..isc-1 / iec+1..ied strips of the
// compute rows (one thread per halo cell of a layer)
template <typename T>
__device__ __forceinline__ void halo_copy_body(const Dims &d, const T *__restrict__ hin, T *__restrict__ h, int bx,
                                               int bz)
{
    const int k = 1 + bz;
    const int rows_out = (d.jsc - d.jsd) + (d.jed - d.jec);
    const int side = (d.isc - d.isd) + (d.ied - d.iec);
    const long long n_rows = (long long)rows_out * d.n1;
    const long long t = (long long)bx * 256 + threadIdx.x;
    int i, j;
    if (t < n_rows) {
        const int r = (int)(t / d.n1);
        i = d.isd + (int)(t % d.n1);
        j = r < d.jsc - d.jsd ? d.jsd + r : d.jec + 1 + (r - (d.jsc - d.jsd));
    } else {
        const long long t2 = t - n_rows;
        if (t2 >= (long long)side * (d.jec - d.jsc + 1)) return;
        const int c = (int)(t2 % side);
        j = d.jsc + (int)(t2 / side);
        i = c < d.isc - d.isd ? d.isd + c : d.iec + 1 + (c - (d.isc - d.isd));
    }
    const long long o = IX3(d, i, j, k);
    h[o] = hin[o];
}

template <typename T>
__global__ void __launch_bounds__(256) k_halo_copy(Dims d, const T *__restrict__ hin, T *__restrict__ h)
{
    halo_copy_body(d, hin, h, blockIdx.x, blockIdx.y);
}

// halo cells of a layer of the grid
inline long long halo_cells(const Dims &d)
{
    return (long long)((d.jsc - d.jsd) + (d.jed - d.jec)) * d.n1 +
           (long long)((d.isc - d.isd) + (d.ied - d.iec)) * (d.jec - d.jsc + 1);
}

// The meridional counterpart of continuity_zonal_convergence:
// h(i, j) = max(hin - dt*IareaT*(vh(i, J=j) - vh(i, J=j-1)), 0) on the compute domain.
template <typename T>
__device__ __forceinline__ void conv_y_body(const Dims &d, const T *hin, const T *__restrict__ vh, double dt,
                                            const double *__restrict__ IareaT, T *h, int bx, int by, int bz)
{
    const int i = d.isc + bx * 128 + threadIdx.x;
    const int j = d.jsc + by;
    const int k = 1 + bz;
    if (i > d.iec) return;
    const long long c = IX3(d, i, j, k);
    const T dtIareaT = (T)(dt * IareaT[IX2(d, i, j)]);
    h[c] = vmax(hin[c] - dtIareaT * (vh[c] - vh[c - d.n1]), T(0.0));
}

template <typename T>
__global__ void __launch_bounds__(128) k_convergence_y(Dims d, const T *hin, const T *__restrict__ vh, double dt,
                                                       const double *__restrict__ IareaT, T *h)
{
    conv_y_body(d, hin, vh, dt, IareaT, h, blockIdx.x, blockIdx.y, blockIdx.z);
}

// the ensemble forms: one launch over every member (blocks flattened x, y, k)
template <typename T>
__global__ void __launch_bounds__(256) k_halo_copy_batch(const ContArgsT<T> *__restrict__ args, BatchMap mp)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const ContArgsT<T> &a = args[m];
    halo_copy_body(a.d, a.hin, a.h, local % mp.nbx[m], local / mp.nbx[m]);
}

template <typename T>
__global__ void __launch_bounds__(128) k_convergence_y_batch(const ContArgsT<T> *__restrict__ args, BatchMap mp)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const ContArgsT<T> &a = args[m];
    const int r = local / mp.nbx[m];
    conv_y_body(a.d, a.hin, a.uh, a.dt, a.IareaT, a.h, local % mp.nbx[m], r % mp.nby[m], r / mp.nby[m]);
}

// continuity_zonal_convergence (:148-176) with hin; optionally copies hin into
// the halo of h (the driver's `h = h0` reset folded into the predictor call).
__global__ void k_convergence(Dims d, const double *hin, const double *__restrict__ uh, double dt,
                              const double *__restrict__ IareaT, double *h, int halo_copy)
{
    const int i = d.isd + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsd + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.ied) return;
    const long long c = IX3(d, i, j, k);
    const bool inside = i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec;
    if (inside) {
        const double v = hin[c] - dt * IareaT[IX2(d, i, j)] * (uh[c] - uh[c - 1]);
        h[c] = dmax(v, 0.0);
    } else if (halo_copy) {
        h[c] = hin[c];
    }
}

// FP64 block shape beyond 32 levels: false = 64 faces x 12 waves of 7 levels
// (1.21 ms per call at 720x720x75), true = 32 faces x 8 waves x 2 groups of 5
// levels, two blocks per CU (1.31 ms: 128 VGPRs with spills, and a block's
// latency chains -- the column folds, the barriers, the loads -- do not
// shorten with fewer faces)
constexpr bool kFP64Narrow = false;

// Blocks in x of a pass with FW faces per block (x: blocks overlap by two faces)
template <bool Y>
unsigned blk_nbx(const Dims &d, int FW)
{
    return Y ? (unsigned)((d.iec - d.isc + 1 + FW - 1) / FW) : (unsigned)((d.iec - d.isc + 1 + FW - 3) / (FW - 2));
}

// A block shape of CW waves of K level slots (64 faces): levels spread evenly
// over the waves when CW*(K-1) <= nk (every shape of 1 or 2 slots, 8 x 4 from
// 24 levels, 12 x 7 and 8 x 10 from 72)
template <typename T, bool Y, int CW, int K>
void launch_shape(const ContArgsT<T> &a, int mono, dim3 g, hipStream_t s)
{
    const dim3 b(64 * CW);
    constexpr size_t lds = sizeof(ContShared<T, CW, K>);
    if constexpr (K <= 2) {
        k_continuity_blk<T, Y, CW, K, CF, true><<<g, b, lds, s>>>(a, mono);
    } else {
        if (a.d.nk >= CW * (K - 1)) k_continuity_blk<T, Y, CW, K, CF, true><<<g, b, lds, s>>>(a, mono);
        else k_continuity_blk<T, Y, CW, K, CF, false><<<g, b, lds, s>>>(a, mono);
    }
}

// One pass over nby rows of faces; the block shape follows nk and the type
template <typename T, bool Y>
int launch_blk(const ContArgsT<T> &a, int mono, unsigned nby, hipStream_t s)
{
    const int nk = a.d.nk;
    const dim3 g(blk_nbx<Y>(a.d, CF), nby);
    if (nk <= 8) launch_shape<T, Y, 8, 1>(a, mono, g, s);
    else if (nk <= 16) launch_shape<T, Y, 8, 2>(a, mono, g, s);
    else if (nk <= 32) launch_shape<T, Y, 8, 4>(a, mono, g, s);
    // FP64 beyond 32 levels: 12 waves of 7 levels (three waves per SIMD); the
    // FP32 form: 8 waves of 10 levels (two blocks per CU)
    else if (sizeof(T) == 8 && !kFP64Narrow) launch_shape<T, Y, 12, 7>(a, mono, g, s);
    // the alternative: blocks of 32 faces, 8 waves x 2 level groups of 5
    // levels (64 KB of LDS, two blocks per CU)
    else if (sizeof(T) == 8)
        k_continuity_blk<T, Y, 8, 5, 32><<<dim3(blk_nbx<Y>(a.d, 32), nby), dim3(64 * 8), sizeof(ContShared<T, 8, 5, 32>),
                                            s>>>(a, mono);
    else launch_shape<T, Y, 8, 10>(a, mono, g, s);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

// Raises the dynamic LDS limit of every block-kernel instantiation of type T
// that fits the device
template <typename T>
int set_blk_lds(long long max_lds)
{
    auto raise = [max_lds](const void *kern, long long bytes) -> hipError_t {
        if (bytes > max_lds) return hipSuccess;
        return hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    };
#define MOM6H_BLK_LDS(Y, CW, K)                                                                                    \
    MOM6H_CHECK(raise((const void *)k_continuity_blk<T, Y, CW, K, CF, true>, sizeof(ContShared<T, CW, K>)));        \
    MOM6H_CHECK(raise((const void *)k_continuity_blk<T, Y, CW, K, CF, false>, sizeof(ContShared<T, CW, K>)));       \
    MOM6H_CHECK(raise((const void *)k_continuity_batch<T, Y, CW, K, true>, sizeof(ContShared<T, CW, K>)));
#define MOM6H_BATCH_PLAIN_LDS(Y, K)                                                                                \
    MOM6H_CHECK(raise((const void *)k_continuity_batch<T, Y, 8, K, false>, sizeof(ContShared<T, 8, K>)));
    MOM6H_BLK_LDS(false, 8, 1) MOM6H_BLK_LDS(false, 8, 2) MOM6H_BLK_LDS(false, 8, 4) MOM6H_BLK_LDS(false, 8, 10)
    MOM6H_BLK_LDS(true, 8, 1) MOM6H_BLK_LDS(true, 8, 2) MOM6H_BLK_LDS(true, 8, 4) MOM6H_BLK_LDS(true, 8, 10)
    MOM6H_BATCH_PLAIN_LDS(false, 4) MOM6H_BATCH_PLAIN_LDS(false, 7) MOM6H_BATCH_PLAIN_LDS(false, 10)
    MOM6H_BATCH_PLAIN_LDS(true, 4) MOM6H_BATCH_PLAIN_LDS(true, 7) MOM6H_BATCH_PLAIN_LDS(true, 10)
    MOM6H_CHECK(raise((const void *)k_continuity_batch<T, false, 8, 7, true>, sizeof(ContShared<T, 8, 7>)));
    MOM6H_CHECK(raise((const void *)k_continuity_batch<T, true, 8, 7, true>, sizeof(ContShared<T, 8, 7>)));
#undef MOM6H_BATCH_PLAIN_LDS
#undef MOM6H_BLK_LDS
    for (const bool Y : {false, true}) {
        const void *even = Y ? (const void *)k_continuity_blk<T, true, 12, 7, CF, true>
                             : (const void *)k_continuity_blk<T, false, 12, 7, CF, true>;
        const void *plain = Y ? (const void *)k_continuity_blk<T, true, 12, 7, CF, false>
                              : (const void *)k_continuity_blk<T, false, 12, 7, CF, false>;
        const void *narrow = Y ? (const void *)k_continuity_blk<T, true, 8, 5, 32>
                               : (const void *)k_continuity_blk<T, false, 8, 5, 32>;
        MOM6H_CHECK(raise(even, sizeof(ContShared<T, 12, 7>)));
        MOM6H_CHECK(raise(plain, sizeof(ContShared<T, 12, 7>)));
        MOM6H_CHECK(raise(narrow, sizeof(ContShared<T, 8, 5, 32>)));
    }
    return MOM6H_OK;
}

template <typename T>
void launch_halo_copy(const Dims &d, const T *hin, T *h, hipStream_t s)
{
    const dim3 bh(256), gh((unsigned)((halo_cells(d) + 255) / 256), d.nk);
    k_halo_copy<T><<<gh, bh, 0, s>>>(d, hin, h);
}

bool blk_fits(const mom6h_continuity_cs *cs)
{
    // the largest FP64 blocks: 12 x 7 levels (one grid), 8 x 10 (ensembles)
    constexpr long long bytes = sizeof(ContShared<double, 12, 7>) > sizeof(ContShared<double, 8, 10>)
                                    ? sizeof(ContShared<double, 12, 7>)
                                    : sizeof(ContShared<double, 8, 10>);
    return cs->d.nk <= 80 && !cs->upwind_1st && !cs->simple_2nd && bytes <= cs->max_lds;
}

// The zonal pass of the block kernel for either type
template <typename T>
int continuity_blk_x(mom6h_continuity_cs *cs, const T *u, const T *hin, T *h, T *uh, double dt, const double *uhbt,
                     const mom6h_bt_cont *bt, T *bt_h_u, T *u_cor, double *du_cor, int hin_halo_to_h, hipStream_t s)
{
    const Dims &d = cs->d;
    ContArgsT<T> a{d, u, hin, h, uh, dt, uhbt, cs->IareaT, cs->IdxT, cs->dy_Cu, cs->dxT, cs->dxCu,
                   cs->mask2dCu, cs->mask2dT, cs->tol_eta, cs->tol_vel, cs->CFL_limit_adjust, 2.0 * 1.0e-10,
                   u_cor, du_cor,
                   bt ? bt->FA_u_EE : nullptr, bt ? bt->FA_u_E0 : nullptr, bt ? bt->FA_u_W0 : nullptr,
                   bt ? bt->FA_u_WW : nullptr, bt ? bt->uBT_WW : nullptr, bt ? bt->uBT_EE : nullptr,
                   bt ? bt_h_u : nullptr};
    int rc = launch_blk<T, false>(a, cs->monotonic, (unsigned)(d.jec - d.jsc + 1), s);
    if (rc) return rc;
    if (hin_halo_to_h) {
        launch_halo_copy(d, hin, h, s);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

// The meridional pass (faces and convergence) for either type
template <typename T>
int continuity_blk_y(mom6h_continuity_cs *cs, const T *v, const T *hin, T *h, T *vh, double dt, const double *vhbt,
                     const mom6h_bt_cont *bt, T *bt_h_v, T *v_cor, double *dv_cor, int hin_halo_to_h, hipStream_t s)
{
    const Dims &d = cs->d;
    ContArgsT<T> a{d, v, hin, h, vh, dt, vhbt, cs->IareaT, cs->IdyT, cs->dx_Cv, cs->dyT, cs->dyCv, cs->mask2dCv,
                   cs->mask2dT, cs->tol_eta, cs->tol_vel, cs->CFL_limit_adjust, 2.0 * 1.0e-10, v_cor, dv_cor,
                   bt ? bt->FA_u_EE : nullptr, bt ? bt->FA_u_E0 : nullptr, bt ? bt->FA_u_W0 : nullptr,
                   bt ? bt->FA_u_WW : nullptr, bt ? bt->uBT_WW : nullptr, bt ? bt->uBT_EE : nullptr,
                   bt ? bt_h_v : nullptr};
    int rc = launch_blk<T, true>(a, cs->monotonic, (unsigned)(d.jec - d.jsc + 2), s);
    if (rc) return rc;
    {
        dim3 b(128), gc((d.iec - d.isc + 1 + 127) / 128, d.jec - d.jsc + 1, d.nk);
        k_convergence_y<T><<<gc, b, 0, s>>>(d, hin, vh, dt, cs->IareaT, h);
        MOM6H_LAUNCH_CHECK();
    }
    if (hin_halo_to_h) {
        launch_halo_copy(d, hin, h, s);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

// One pass (zonal, or meridional with its convergence) of every member of an
// ensemble: one launch per kernel.  h_args mirrors d_args on the host.
template <typename T>
int continuity_batch(const ContArgsT<T> *d_args, const ContArgsT<T> *h_args, int n, bool y_pass, bool halo_copy,
                     long long max_lds, hipStream_t s)
{
    MOM6H_REQUIRE(n >= 1 && n <= kMaxMembers, "continuity batch: 1..%d members", kMaxMembers);
    BatchMap mp{};
    mp.n = n;
    int nk_max = 1;
    for (int m = 0; m < n; ++m) {
        const Dims &d = h_args[m].d;
        mp.nbx[m] = (int)(y_pass ? blk_nbx<true>(d, CF) : blk_nbx<false>(d, CF));
        mp.nby[m] = y_pass ? d.jec - d.jsc + 2 : d.jec - d.jsc + 1;
        mp.first[m + 1] = mp.first[m] + mp.nbx[m] * mp.nby[m];
        nk_max = d.nk > nk_max ? d.nk : nk_max;
    }
    MOM6H_REQUIRE(nk_max <= 80 && (long long)sizeof(ContShared<T, 8, 10>) <= max_lds, "continuity batch: nk <= 80");
    // one launch per block shape: 8 waves of K level slots, K from the
    // member's depth (a member of 25 levels takes 8 x 4 = 32 slots rather than
    // the 80 of the deepest member's 8 x 10); the levels spread evenly when
    // every member of the launch has nk >= 8*(K-1)
    const int shapes[5] = {1, 2, 4, 7, 10};
    auto shape_of = [](int nk) { return nk <= 8 ? 1 : nk <= 16 ? 2 : nk <= 32 ? 4 : nk <= 56 ? 7 : 10; };
    for (const int K : shapes) {
        BatchMap g{};
        MemberIndex mi{};
        bool even = true;
        for (int m = 0; m < n; ++m) {
            const int nk = h_args[m].d.nk;
            if (shape_of(nk) != K) continue;
            const int q = g.n++;
            mi.v[q] = m;
            g.nbx[q] = mp.nbx[m];
            g.nby[q] = mp.nby[m];
            g.first[q + 1] = g.first[q] + g.nbx[q] * g.nby[q];
            even = even && nk >= 8 * (K - 1);
        }
        if (g.n == 0) continue;
        const unsigned nb = (unsigned)g.first[g.n];
        const dim3 b(64 * 8);
#define CB(Y, KK)                                                                                             \
    if (even) k_continuity_batch<T, Y, 8, KK, true><<<nb, b, sizeof(ContShared<T, 8, KK>), s>>>(d_args, g, mi); \
    else k_continuity_batch<T, Y, 8, KK, false><<<nb, b, sizeof(ContShared<T, 8, KK>), s>>>(d_args, g, mi);
#define CBY(KK) \
    if (y_pass) { CB(true, KK) } else { CB(false, KK) }
        if (K == 1) {   // shapes of 1 and 2 slots always spread evenly
            if (y_pass) k_continuity_batch<T, true, 8, 1, true><<<nb, b, sizeof(ContShared<T, 8, 1>), s>>>(d_args, g, mi);
            else k_continuity_batch<T, false, 8, 1, true><<<nb, b, sizeof(ContShared<T, 8, 1>), s>>>(d_args, g, mi);
        } else if (K == 2) {
            if (y_pass) k_continuity_batch<T, true, 8, 2, true><<<nb, b, sizeof(ContShared<T, 8, 2>), s>>>(d_args, g, mi);
            else k_continuity_batch<T, false, 8, 2, true><<<nb, b, sizeof(ContShared<T, 8, 2>), s>>>(d_args, g, mi);
        } else if (K == 4) {
            CBY(4)
        } else if (K == 7) {
            CBY(7)
        } else {
            CBY(10)
        }
#undef CBY
#undef CB
        MOM6H_LAUNCH_CHECK();
    }
    if (y_pass) {
        BatchMap mc{};
        mc.n = n;
        for (int m = 0; m < n; ++m) {
            const Dims &d = h_args[m].d;
            mc.nbx[m] = (d.iec - d.isc + 1 + 127) / 128;
            mc.nby[m] = d.jec - d.jsc + 1;
            mc.first[m + 1] = mc.first[m] + mc.nbx[m] * mc.nby[m] * d.nk;
        }
        k_convergence_y_batch<T><<<(unsigned)mc.first[n], 128, 0, s>>>(d_args, mc);
        MOM6H_LAUNCH_CHECK();
    }
    if (halo_copy) {
        BatchMap mh{};
        mh.n = n;
        for (int m = 0; m < n; ++m) {
            const Dims &d = h_args[m].d;
            mh.nbx[m] = (int)((halo_cells(d) + 255) / 256);
            mh.nby[m] = d.nk;
            mh.first[m + 1] = mh.first[m] + mh.nbx[m] * d.nk;
        }
        k_halo_copy_batch<T><<<(unsigned)mh.first[n], 256, 0, s>>>(d_args, mh);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

template int continuity_batch<float>(const ContArgsT<float> *, const ContArgsT<float> *, int, bool, bool, long long,
                                     hipStream_t);
template int continuity_batch<double>(const ContArgsT<double> *, const ContArgsT<double> *, int, bool, bool, long long,
                                      hipStream_t);

// The arguments the single-grid block kernels receive (for the ensemble tables)
template <typename T>
ContArgsT<T> continuity_args(const mom6h_continuity_cs *cs, bool y_pass, const T *u, const T *hin, T *h, T *uh,
                             double dt, const mom6h_bt_cont *bt, T *bt_h_u, T *u_cor, double *du_cor)
{
    ContArgsT<T> a{};
    a.d = cs->d; a.u = u; a.hin = hin; a.h = h; a.uh = uh; a.dt = dt; a.uhbt = nullptr;
    a.IareaT = cs->IareaT;
    a.IdxT = y_pass ? cs->IdyT : cs->IdxT;
    a.dy_Cu = y_pass ? cs->dx_Cv : cs->dy_Cu;
    a.dxT = y_pass ? cs->dyT : cs->dxT;
    a.dxCu = y_pass ? cs->dyCv : cs->dxCu;
    a.mask2dCu = y_pass ? cs->mask2dCv : cs->mask2dCu;
    a.mask2dT = cs->mask2dT;
    a.tol_eta0 = cs->tol_eta; a.tol_vel = cs->tol_vel; a.CFL_limit_adjust = cs->CFL_limit_adjust;
    a.h_min = 2.0 * 1.0e-10;
    a.u_cor = u_cor; a.du_cor = du_cor;
    if (bt) {
        a.FA_u_EE = bt->FA_u_EE; a.FA_u_E0 = bt->FA_u_E0; a.FA_u_W0 = bt->FA_u_W0; a.FA_u_WW = bt->FA_u_WW;
        a.uBT_WW = bt->uBT_WW; a.uBT_EE = bt->uBT_EE; a.bt_h_u = bt_h_u;
    }
    a.monotonic = cs->monotonic;
    return a;
}

template ContArgsT<float> continuity_args<float>(const mom6h_continuity_cs *, bool, const float *, const float *,
                                                 float *, float *, double, const mom6h_bt_cont *, float *, float *,
                                                 double *);
template ContArgsT<double> continuity_args<double>(const mom6h_continuity_cs *, bool, const double *, const double *,
                                                   double *, double *, double, const mom6h_bt_cont *, double *,
                                                   double *, double *);

int continuity_f32(mom6h_continuity_cs *cs, const float *u, const float *hin, float *h, float *uh, double dt,
                   const mom6h_bt_cont *bt, float *bt_h_u, float *u_cor, double *du_cor, int hin_halo_to_h,
                   hipStream_t s)
{
    MOM6H_REQUIRE(blk_fits(cs) && (long long)sizeof(ContShared<float, 8, 10>) <= cs->max_lds,
                  "continuity (FP32 state): needs nk <= 80 and the PPM reconstruction");
    return continuity_blk_x<float>(cs, u, hin, h, uh, dt, nullptr, bt, bt_h_u, u_cor, du_cor, hin_halo_to_h, s);
}

int continuity_y_f32(mom6h_continuity_cs *cs, const float *v, const float *hin, float *h, float *vh, double dt,
                     const mom6h_bt_cont *bt, float *bt_h_v, float *v_cor, double *dv_cor, int hin_halo_to_h,
                     hipStream_t s)
{
    MOM6H_REQUIRE(cs->IdyT != nullptr, "continuity (FP32 state): meridional metrics not set");
    MOM6H_REQUIRE(blk_fits(cs), "continuity (FP32 state): needs nk <= 80 and the PPM reconstruction");
    return continuity_blk_y<float>(cs, v, hin, h, vh, dt, nullptr, bt, bt_h_v, v_cor, dv_cor, hin_halo_to_h, s);
}

#define RC_(expr)              \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_continuity_init(mom6h_continuity_cs **out, int isd, int ied, int jsd, int jed, int isc,
                          int iec, int jsc, int jec, int nk, const double *IareaT,
                          const double *IdxT, const double *dy_Cu, const double *mask2dT,
                          int monotonic, const double *dxT, const double *dxCu,
                          const double *mask2dCu)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_continuity_init: NULL cs");
    MOM6H_REQUIRE(isd < isc - 2 && iec + 2 < ied && jsd <= jsc && jec <= jed && isc <= iec &&
                      jsc <= jec && nk >= 1,
                  "mom6h_continuity_init: PPM needs 3 halo columns (isd..isc-3, iec+3..ied)");
    MOM6H_REQUIRE(IareaT && IdxT && dy_Cu && mask2dT && dxT && dxCu && mask2dCu,
                  "mom6h_continuity_init: NULL metric");
    auto *cs = new mom6h_continuity_cs();
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    // continuity_init (mom6_continuity_serial.F90:66-76)
    cs->monotonic = monotonic;
    cs->upwind_1st = 0;
    cs->simple_2nd = 0;
    cs->tol_eta = 1.0e-12;
    cs->tol_vel = 3.0e8;
    cs->CFL_limit_adjust = 0.5;
    {
        // the block kernels keep a whole column block in LDS (up to 125 KB)
        int dev = 0, lds = 0;
        MOM6H_CHECK(hipGetDevice(&dev));
        MOM6H_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cs->max_lds = lds;
        int rc0 = set_blk_lds<double>(lds);
        rc0 = rc0 ? rc0 : set_blk_lds<float>(lds);
        if (rc0) {
            delete cs;
            return rc0;
        }
    }
    const size_t n2 = (size_t)cs->d.s2, n3 = n2 * nk;
    int rc = 0;
    rc = rc ? rc : dev_upload(&cs->IareaT, IareaT, n2);
    rc = rc ? rc : dev_upload(&cs->IdxT, IdxT, n2);
    rc = rc ? rc : dev_upload(&cs->dy_Cu, dy_Cu, n2);
    rc = rc ? rc : dev_upload(&cs->mask2dT, mask2dT, n2);
    rc = rc ? rc : dev_upload(&cs->dxT, dxT, n2);
    rc = rc ? rc : dev_upload(&cs->dxCu, dxCu, n2);
    rc = rc ? rc : dev_upload(&cs->mask2dCu, mask2dCu, n2);
    rc = rc ? rc : dev_alloc(&cs->hW, n3, true);
    rc = rc ? rc : dev_alloc(&cs->hE, n3, true);
    if (rc) {
        mom6h_continuity_end(cs);
        return rc;
    }
    *out = cs;
    return MOM6H_OK;
}

int mom6h_continuity_ppm(mom6h_continuity_cs *cs, const double *u, const double *hin, double *h,
                         double *uh, double dt, const double *uhbt, const mom6h_bt_cont *bt,
                         double *u_cor, double *du_cor, int hin_halo_to_h, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_continuity_ppm: NULL cs");
    MOM6H_REQUIRE(u && hin && h && uh, "mom6h_continuity_ppm: NULL array");
    MOM6H_REQUIRE((const void *)hin != (const void *)h, "mom6h_continuity_ppm: hin and h must be distinct arrays");
    MOM6H_REQUIRE(dt > 0.0, "mom6h_continuity_ppm: dt must be positive");
    MOM6H_REQUIRE(!bt || (bt->FA_u_EE && bt->FA_u_E0 && bt->FA_u_W0 && bt->FA_u_WW && bt->uBT_WW &&
                          bt->uBT_EE && bt->h_u),
                  "mom6h_continuity_ppm: incomplete BT_cont");
    const Dims &d = cs->d;
    hipStream_t s = (hipStream_t)stream;
    if (blk_fits(cs))
        return continuity_blk_x<double>(cs, u, hin, h, uh, dt, uhbt, bt, bt ? bt->h_u : nullptr, u_cor, du_cor,
                                        hin_halo_to_h, s);
    {
        const int ni = (d.iec + 1) - (d.isc - 1) + 1;
        dim3 b(64), g((ni + 63) / 64, d.jec - d.jsc + 1, d.nk);
        k_ppm_edges<<<g, b, 0, s>>>(d, hin, cs->mask2dT, cs->hW, cs->hE, cs->upwind_1st,
                                    cs->simple_2nd, cs->monotonic, 2.0 * 1.0e-10);
        MOM6H_LAUNCH_CHECK();
    }
    {
        const int ni = d.iec - (d.isc - 1) + 1;
        dim3 b(64), g((ni + 63) / 64, d.jec - d.jsc + 1);
        k_mass_flux<<<g, b, 0, s>>>(d, u, hin, cs->hW, cs->hE, uh, dt, uhbt, cs->IareaT, cs->IdxT,
                                    cs->dy_Cu, cs->dxT, cs->dxCu, cs->mask2dCu, cs->tol_eta,
                                    cs->tol_vel, cs->CFL_limit_adjust, u_cor, du_cor,
                                    bt ? bt->FA_u_EE : nullptr, bt ? bt->FA_u_E0 : nullptr,
                                    bt ? bt->FA_u_W0 : nullptr, bt ? bt->FA_u_WW : nullptr,
                                    bt ? bt->uBT_WW : nullptr, bt ? bt->uBT_EE : nullptr,
                                    bt ? bt->h_u : nullptr, bt ? 1 : 0);
        MOM6H_LAUNCH_CHECK();
    }
    {
        dim3 b(128), g((d.n1 + 127) / 128, d.n2, d.nk);
        k_convergence<<<g, b, 0, s>>>(d, hin, uh, dt, cs->IareaT, h, hin_halo_to_h);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

int mom6h_continuity_set_meridional(mom6h_continuity_cs *cs, const double *IdyT, const double *dx_Cv,
                                    const double *dyT, const double *dyCv, const double *mask2dCv)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_continuity_set_meridional: NULL cs");
    MOM6H_REQUIRE(IdyT && dx_Cv && dyT && dyCv && mask2dCv, "mom6h_continuity_set_meridional: NULL metric");
    const Dims &d = cs->d;
    MOM6H_REQUIRE(d.jsd < d.jsc - 2 && d.jec + 2 < d.jed,
                  "mom6h_continuity_set_meridional: PPM needs 3 halo rows (jsd..jsc-3, jec+3..jed)");
    const size_t n2 = (size_t)d.s2;
    for (double *p : {cs->IdyT, cs->dx_Cv, cs->dyT, cs->dyCv, cs->mask2dCv}) dev_free(p);
    cs->IdyT = cs->dx_Cv = cs->dyT = cs->dyCv = cs->mask2dCv = nullptr;
    int rc = 0;
    rc = rc ? rc : dev_upload(&cs->IdyT, IdyT, n2);
    rc = rc ? rc : dev_upload(&cs->dx_Cv, dx_Cv, n2);
    rc = rc ? rc : dev_upload(&cs->dyT, dyT, n2);
    rc = rc ? rc : dev_upload(&cs->dyCv, dyCv, n2);
    rc = rc ? rc : dev_upload(&cs->mask2dCv, mask2dCv, n2);
    return rc;
}

int mom6h_continuity_ppm_y(mom6h_continuity_cs *cs, const double *v, const double *hin, double *h,
                           double *vh, double dt, const double *vhbt, const mom6h_bt_cont *bt,
                           double *v_cor, double *dv_cor, int hin_halo_to_h, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_continuity_ppm_y: NULL cs");
    MOM6H_REQUIRE(cs->IdyT != nullptr, "mom6h_continuity_ppm_y: call mom6h_continuity_set_meridional first");
    MOM6H_REQUIRE(v && hin && h && vh, "mom6h_continuity_ppm_y: NULL array");
    MOM6H_REQUIRE((const void *)hin != (const void *)h, "mom6h_continuity_ppm_y: hin and h must be distinct arrays");
    MOM6H_REQUIRE(dt > 0.0, "mom6h_continuity_ppm_y: dt must be positive");
    MOM6H_REQUIRE(!bt || (bt->FA_u_EE && bt->FA_u_E0 && bt->FA_u_W0 && bt->FA_u_WW && bt->uBT_WW &&
                          bt->uBT_EE && bt->h_u),
                  "mom6h_continuity_ppm_y: incomplete BT_cont");
    MOM6H_REQUIRE(blk_fits(cs),
                  "mom6h_continuity_ppm_y: the meridional pass supports nk <= 80 with the PPM reconstruction");
    return continuity_blk_y<double>(cs, v, hin, h, vh, dt, vhbt, bt, bt ? bt->h_u : nullptr, v_cor, dv_cor,
                                    hin_halo_to_h, (hipStream_t)stream);
}

int mom6h_continuity_end(mom6h_continuity_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->IareaT, cs->IdxT, cs->dy_Cu, cs->mask2dT, cs->dxT, cs->dxCu,
                      cs->mask2dCu, cs->hW, cs->hE, cs->IdyT, cs->dx_Cv, cs->dyT, cs->dyCv, cs->mask2dCv})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

int mom6h_bt_cont_alloc(mom6h_bt_cont *bt, int isd, int ied, int jsd, int jed, int nk)
{
    MOM6H_REQUIRE(bt != nullptr && ied >= isd && jed >= jsd && nk >= 1, "mom6h_bt_cont_alloc: bad args");
    *bt = mom6h_bt_cont{};
    const size_t n2 = (size_t)(ied - isd + 1) * (jed - jsd + 1);
    int rc = 0;
    for (double **p : {&bt->FA_u_EE, &bt->FA_u_E0, &bt->FA_u_W0, &bt->FA_u_WW, &bt->uBT_WW, &bt->uBT_EE})
        rc = rc ? rc : dev_alloc(p, n2, true);
    rc = rc ? rc : dev_alloc(&bt->h_u, n2 * nk, true);
    if (rc) mom6h_bt_cont_free(bt);
    return rc;
}

int mom6h_bt_cont_free(mom6h_bt_cont *bt)
{
    if (!bt) return MOM6H_OK;
    for (double *p : {bt->FA_u_EE, bt->FA_u_E0, bt->FA_u_W0, bt->FA_u_WW, bt->uBT_WW, bt->uBT_EE, bt->h_u})
        dev_free(p);
    *bt = mom6h_bt_cont{};
    return MOM6H_OK;
}

}  // extern "C"
