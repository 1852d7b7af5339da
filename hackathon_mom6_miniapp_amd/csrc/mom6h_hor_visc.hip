// mom6h_hor_visc.hip -- horizontal viscosity on gfx950.
//
// Arithmetic of src/serial/mom6_hor_visc_serial.F90:464-1109 with the
// Laplacian (constant Kh) and biharmonic (constant Ah) operators, free- or
// no-slip; and (k_hv_closures, further below) the Smagorinsky and Leith
// viscosities with the thickness-aware bounds.  One 3D launch computes every layer: each thread forms the
// strain/stress values its diffu/diffv stencil needs directly from u, v, h
// (no 2D work arrays, no per-layer launches).  On grids whose metric arrays
// are constant the constant-coefficient operator runs as a tiled layer walk
// with LDS staging instead (k_hv_tile), bit-identical to the general kernel.
//
// The serial code keeps ONE 2D array for Kh (and Ah) at h- and q-points: the
// h-point pass writes is..ie x js..je but the stress uses is-1..ie+1 x
// js-1..je+1 (:744-819), so the ring takes what the previous layer's q-point
// pass left there (Kh_bg_xy on is-1..ie x js-1..je, zero elsewhere, and zero
// everywhere on the first layer of the first call).  Likewise Del2u/Del2v are
// only formed on their own ranges (:587-607) and read as zero outside.  Those
// values only reach diffu on column is-1 and ie and diffv on rows js-1 and je
// (biharmonic: also the first interior column/row), and are reproduced here.
#include "mom6h_internal.h"

#include <initializer_list>
#include <utility>

namespace mom6h {

struct HV {
    Dims d;
    double Kh, Ah, h_neglect;
    int laplacian, biharmonic, no_slip, first_call;
    const double *DY_dxT, *DX_dyT, *DY_dxBu, *DX_dyBu, *IdyCu, *IdxCu, *IdyCv, *IdxCv;
    const double *IareaCu, *IareaCv, *mask2dT, *mask2dBu, *reduction_xx, *reduction_xy;
    const double *dy2h, *dx2h, *dy2q, *dx2q;
    const double *Idx2dyCu, *Idxdy2u, *Idx2dyCv, *Idxdy2v;
    double *Del2u, *Del2v;
    // closures (the general kernel): Smagorinsky_Kh/_Ah, Leith_Kh/_Ah,
    // better_bound_Kh/_Ah = bound_Kh/_Ah (hor_visc_init :179-193)
    int smag_Kh, smag_Ah, leith_Kh, leith_Ah, bound;
    const double *areaT, *dxBu, *dyBu, *IdxBu, *IdyBu;
    double *ring_Kh, *ring_Ah;
};

// horizontal tension at an h point (:503-530)
__device__ inline double sh_xx_at(const HV &p, const double *u, const double *v, int i, int j, int k)
{
    const Dims &d = p.d;
    const double dudx = p.DY_dxT[IX2(d, i, j)] *
                        ((p.IdyCu[IX2(d, i, j)] * u[IX3(d, i, j, k)]) - (p.IdyCu[IX2(d, i - 1, j)] * u[IX3(d, i - 1, j, k)]));
    const double dvdy = p.DX_dyT[IX2(d, i, j)] *
                        ((p.IdxCv[IX2(d, i, j)] * v[IX3(d, i, j, k)]) - (p.IdxCv[IX2(d, i, j - 1)] * v[IX3(d, i, j - 1, k)]));
    return dudx - dvdy;
}

// shearing strain at a q point with the slip condition (:513-547)
__device__ inline double sh_xy_at(const HV &p, const double *u, const double *v, int I, int J, int k)
{
    const Dims &d = p.d;
    const double dvdx = p.DY_dxBu[IX2(d, I, J)] *
                        ((v[IX3(d, I + 1, J, k)] * p.IdyCv[IX2(d, I + 1, J)]) - (v[IX3(d, I, J, k)] * p.IdyCv[IX2(d, I, J)]));
    const double dudy = p.DX_dyBu[IX2(d, I, J)] *
                        ((u[IX3(d, I, J + 1, k)] * p.IdxCu[IX2(d, I, J + 1)]) - (u[IX3(d, I, J, k)] * p.IdxCu[IX2(d, I, J)]));
    const double m = p.mask2dBu[IX2(d, I, J)];
    return (p.no_slip ? (2.0 - m) : m) * (dvdx + dudy);
}

// Del2u / Del2v (:587-607) on their ranges, zero elsewhere (never written there).
__global__ void k_hv_del2(HV p, const double *__restrict__ u, const double *__restrict__ v)
{
    const Dims &d = p.d;
    const int i = d.isc - 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc - 1 + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.iec) return;
    const long long c = IX3(d, i, j, k), c2 = IX2(d, i, j);
    if (j >= d.jsc) {
        p.Del2u[c] = p.Idx2dyCu[c2] * ((p.dx2q[c2] * sh_xy_at(p, u, v, i, j, k)) -
                                       (p.dx2q[IX2(d, i, j - 1)] * sh_xy_at(p, u, v, i, j - 1, k))) +
                     p.Idxdy2u[c2] * ((p.dy2h[IX2(d, i + 1, j)] * sh_xx_at(p, u, v, i + 1, j, k)) -
                                      (p.dy2h[c2] * sh_xx_at(p, u, v, i, j, k)));
    }
    if (i >= d.isc) {
        p.Del2v[c] = p.Idxdy2v[c2] * ((p.dy2q[c2] * sh_xy_at(p, u, v, i, j, k)) -
                                      (p.dy2q[IX2(d, i - 1, j)] * sh_xy_at(p, u, v, i - 1, j, k))) -
                     p.Idx2dyCv[c2] * ((p.dx2h[IX2(d, i, j + 1)] * sh_xx_at(p, u, v, i, j + 1, k)) -
                                       (p.dx2h[c2] * sh_xx_at(p, u, v, i, j, k)));
    }
}

// Kh or Ah seen at an h point of the stress ring (see the file comment).
__device__ inline double visc_h(const HV &p, double val, int i, int j, int k)
{
    const Dims &d = p.d;
    if (i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec) return val;
    if (i >= d.isc - 1 && i <= d.iec && j >= d.jsc - 1 && j <= d.jec)
        return (k == 1 && p.first_call) ? 0.0 : val;
    return 0.0;
}

// str_xx * h (:807-895) at h point (i,j)
__device__ inline double str_xx_at(const HV &p, const double *u, const double *v, const double *h,
                                   int i, int j, int k)
{
    const Dims &d = p.d;
    const double sh = sh_xx_at(p, u, v, i, j, k);
    double str = p.laplacian ? -(visc_h(p, p.Kh, i, j, k) * sh) : 0.0;
    if (p.biharmonic) {
        const double d_del2u = (p.IdyCu[IX2(d, i, j)] * p.Del2u[IX3(d, i, j, k)]) -
                               (p.IdyCu[IX2(d, i - 1, j)] * p.Del2u[IX3(d, i - 1, j, k)]);
        const double d_del2v = (p.IdxCv[IX2(d, i, j)] * p.Del2v[IX3(d, i, j, k)]) -
                               (p.IdxCv[IX2(d, i, j - 1)] * p.Del2v[IX3(d, i, j - 1, k)]);
        const double d_str = visc_h(p, p.Ah, i, j, k) *
                             ((p.DY_dxT[IX2(d, i, j)] * d_del2u) - (p.DX_dyT[IX2(d, i, j)] * d_del2v));
        str = str + d_str;
    }
    return str * (h[IX3(d, i, j, k)] * p.reduction_xx[IX2(d, i, j)]);
}

// str_xy * hq (:954-1053) at q point (I,J)
__device__ inline double str_xy_at(const HV &p, const double *u, const double *v, const double *h,
                                   int I, int J, int k)
{
    const Dims &d = p.d;
    const long long c2 = IX2(d, I, J);
    const double sh = sh_xy_at(p, u, v, I, J, k);
    double str = p.laplacian ? -(p.Kh * sh) : 0.0;
    if (p.biharmonic) {
        const double d_str =
            p.Ah * ((p.DY_dxBu[c2] * ((p.Del2v[IX3(d, I + 1, J, k)] * p.IdyCv[IX2(d, I + 1, J)]) -
                                      (p.Del2v[IX3(d, I, J, k)] * p.IdyCv[c2]))) +
                    (p.DX_dyBu[c2] * ((p.Del2u[IX3(d, I, J + 1, k)] * p.IdxCu[IX2(d, I, J + 1)]) -
                                      (p.Del2u[IX3(d, I, J, k)] * p.IdxCu[c2]))));
        str = str + d_str;
    }
    const double hq = 0.25 * ((h[IX3(d, I, J, k)] + h[IX3(d, I + 1, J + 1, k)]) +
                              (h[IX3(d, I + 1, J, k)] + h[IX3(d, I, J + 1, k)]));
    if (p.no_slip) return str * (hq * p.reduction_xy[c2]);
    return str * (hq * p.mask2dBu[c2] * p.reduction_xy[c2]);
}

// diffu, diffv over the full data domain (zero outside their ranges, :486-494, :1059-1078)
__global__ void k_hv_main(HV p, const double *__restrict__ u, const double *__restrict__ v,
                          const double *__restrict__ h, double *__restrict__ diffu,
                          double *__restrict__ diffv)
{
    const Dims &d = p.d;
    const int i = d.isd + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsd + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.ied) return;
    const long long c = IX3(d, i, j, k), c2 = IX2(d, i, j);
    double du = 0.0, dv = 0.0;
    if (i >= d.isc - 1 && i <= d.iec && j >= d.jsc && j <= d.jec) {
        const double h_u = 0.5 * (p.mask2dT[c2] * h[c] + p.mask2dT[IX2(d, i + 1, j)] * h[IX3(d, i + 1, j, k)]);
        du = ((p.IdxCu[c2] * ((p.dx2q[IX2(d, i, j - 1)] * str_xy_at(p, u, v, h, i, j - 1, k)) -
                              (p.dx2q[c2] * str_xy_at(p, u, v, h, i, j, k))) +
               p.IdyCu[c2] * ((p.dy2h[c2] * str_xx_at(p, u, v, h, i, j, k)) -
                              (p.dy2h[IX2(d, i + 1, j)] * str_xx_at(p, u, v, h, i + 1, j, k)))) *
              p.IareaCu[c2]) / (h_u + p.h_neglect);
    }
    if (i >= d.isc && i <= d.iec && j >= d.jsc - 1 && j <= d.jec) {
        const double h_v = 0.5 * (p.mask2dT[c2] * h[c] + p.mask2dT[IX2(d, i, j + 1)] * h[IX3(d, i, j + 1, k)]);
        dv = ((p.IdyCv[c2] * ((p.dy2q[IX2(d, i - 1, j)] * str_xy_at(p, u, v, h, i - 1, j, k)) -
                              (p.dy2q[c2] * str_xy_at(p, u, v, h, i, j, k))) -
               p.IdxCv[c2] * ((p.dx2h[c2] * str_xx_at(p, u, v, h, i, j, k)) -
                              (p.dx2h[IX2(d, i, j + 1)] * str_xx_at(p, u, v, h, i, j + 1, k)))) *
              p.IareaCv[c2]) / (h_v + p.h_neglect);
    }
    diffu[c] = du;
    diffv[c] = dv;
}

// ---------------------------------------------------------------------------
// The constant-coefficient operator on a grid whose metric arrays are constant
// (over is-2..ie+2 x js-2..je+2, every point the stencil reads): a tiled layer
// walk.  A block owns HT_X x TY points of the data domain and walks a range of
// layers.  Per layer it stages u, v, h (and Del2u, Del2v) of the tile and a
// one-cell ring in LDS, forms str_xx*h at every staged h point and str_xy*hq
// at every staged q point once, then diffu and diffv of its points (zero
// outside their ranges).  The masks stay in LDS for the whole walk, and the
// next layer's loads are in flight while the current one is computed.  The
// expressions are k_hv_main's with the constant metric values, so the results
// are bit-identical to it.
constexpr int HT_X = 64, HW = HT_X + 2;

template <bool BIH, int TY>
__global__ void __launch_bounds__(256) k_hv_tile(HV p, HVUniform g, const double *__restrict__ u,
                                                 const double *__restrict__ v, const double *__restrict__ h,
                                                 double *__restrict__ diffu, double *__restrict__ diffv)
{
    constexpr int MH = TY + 2, RPT = TY / 4, NS = (HW * MH + 255) / 256;
    __shared__ double s_u[MH][HW], s_v[MH][HW], s_h[MH][HW], s_mT[MH][HW], s_mBu[MH][HW];
    __shared__ double s_xx[MH][HW], s_xy[MH][HW];
    __shared__ double s_2u[BIH ? MH : 1][BIH ? HW : 1], s_2v[BIH ? MH : 1][BIH ? HW : 1];
    const Dims &d = p.d;
    const int nxy = gridDim.x * gridDim.y;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nxy * gridDim.z);
    const int xy = tile % nxy, kz = tile / nxy;
    const int per = (d.nk + (int)gridDim.z - 1) / (int)gridDim.z;
    const int k0 = 1 + kz * per, k1 = k0 + per - 1 < d.nk ? k0 + per - 1 : d.nk;
    if (k0 > k1) return;
    const int i0 = d.isd + (xy % (int)gridDim.x) * HT_X, j0 = d.jsd + (xy / (int)gridDim.x) * TY;
    const int bi = i0 - 1, bj = j0 - 1;
    const int tid = threadIdx.x, tx = tid % HT_X, ty = tid / HT_X;

    double pu[NS], pv[NS], ph[NS], p2u[BIH ? NS : 1], p2v[BIH ? NS : 1];
    auto load_layer = [&](int k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r, li = t % HW, lj = t / HW, i = bi + li, j = bj + lj;
            pu[r] = 0.0; pv[r] = 0.0; ph[r] = 0.0;
            if (BIH) { p2u[BIH ? r : 0] = 0.0; p2v[BIH ? r : 0] = 0.0; }
            if (t < HW * MH && i >= d.isd && i <= d.ied && j >= d.jsd && j <= d.jed) {
                const long long c = IX3(d, i, j, k);
                pu[r] = u[c]; pv[r] = v[c]; ph[r] = h[c];
                if (BIH) { p2u[BIH ? r : 0] = p.Del2u[c]; p2v[BIH ? r : 0] = p.Del2v[c]; }
            }
        }
    };
    for (int t = tid; t < HW * MH; t += 256) {
        const int li = t % HW, lj = t / HW, i = bi + li, j = bj + lj;
        double mt = 0.0, mb = 0.0;
        if (i >= d.isd && i <= d.ied && j >= d.jsd && j <= d.jed) {
            mt = p.mask2dT[IX2(d, i, j)];
            mb = p.mask2dBu[IX2(d, i, j)];
        }
        s_mT[lj][li] = mt;
        s_mBu[lj][li] = mb;
    }
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    load_layer(k0);
    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            if (t < HW * MH) {
                const int li = t % HW, lj = t / HW;
                s_u[lj][li] = pu[r]; s_v[lj][li] = pv[r]; s_h[lj][li] = ph[r];
                if (BIH) { s_2u[BIH ? lj : 0][BIH ? li : 0] = p2u[BIH ? r : 0]; s_2v[BIH ? lj : 0][BIH ? li : 0] = p2v[BIH ? r : 0]; }
            }
        }
        __syncthreads();
        if (k < k1) load_layer(k + 1);
        // Kh and Ah at the h points of the stress ring: what the previous
        // layer's q-point pass left there, zero on the first layer of the
        // first call (visc_h)
        const bool ring_zero = k == 1 && p.first_call;
        for (int t = tid; t < HW * MH; t += 256) {
            const int li = t % HW, lj = t / HW, i = bi + li, j = bj + lj;
            if (li >= 1 && lj >= 1) {   // str_xx * h at h point (i, j) (:807-895)
                const double dudx = g.DY_dxT * ((g.IdyCu * s_u[lj][li]) - (g.IdyCu * s_u[lj][li - 1]));
                const double dvdy = g.DX_dyT * ((g.IdxCv * s_v[lj][li]) - (g.IdxCv * s_v[lj - 1][li]));
                const double sh = dudx - dvdy;
                const bool in = i >= is && i <= ie && j >= js && j <= je;
                const bool ring = !in && i >= is - 1 && i <= ie && j >= js - 1 && j <= je;
                const bool on = in || (ring && !ring_zero);
                const double Kh = on ? p.Kh : 0.0;
                double str = p.laplacian ? -(Kh * sh) : 0.0;
                if (BIH) {
                    const int qi = BIH ? li : 0, qj = BIH ? lj : 0;
                    const double d_del2u = (g.IdyCu * s_2u[qj][qi]) - (g.IdyCu * s_2u[qj][BIH ? qi - 1 : 0]);
                    const double d_del2v = (g.IdxCv * s_2v[qj][qi]) - (g.IdxCv * s_2v[BIH ? qj - 1 : 0][qi]);
                    const double Ah = on ? p.Ah : 0.0;
                    const double d_str = Ah * ((g.DY_dxT * d_del2u) - (g.DX_dyT * d_del2v));
                    str = str + d_str;
                }
                s_xx[lj][li] = str * (s_h[lj][li] * g.red_xx);
            }
            if (li < HW - 1 && lj < MH - 1) {   // str_xy * hq at q point (i, j) (:954-1053)
                const double dvdx = g.DY_dxBu * ((s_v[lj][li + 1] * g.IdyCv) - (s_v[lj][li] * g.IdyCv));
                const double dudy = g.DX_dyBu * ((s_u[lj + 1][li] * g.IdxCu) - (s_u[lj][li] * g.IdxCu));
                const double m = s_mBu[lj][li];
                const double sh = (p.no_slip ? (2.0 - m) : m) * (dvdx + dudy);
                double str = p.laplacian ? -(p.Kh * sh) : 0.0;
                if (BIH) {
                    const int qi = BIH ? li : 0, qj = BIH ? lj : 0;
                    const double d_str =
                        p.Ah * ((g.DY_dxBu * ((s_2v[qj][BIH ? qi + 1 : 0] * g.IdyCv) - (s_2v[qj][qi] * g.IdyCv))) +
                                (g.DX_dyBu * ((s_2u[BIH ? qj + 1 : 0][qi] * g.IdxCu) - (s_2u[qj][qi] * g.IdxCu))));
                    str = str + d_str;
                }
                const double hq = 0.25 * ((s_h[lj][li] + s_h[lj + 1][li + 1]) + (s_h[lj][li + 1] + s_h[lj + 1][li]));
                s_xy[lj][li] = p.no_slip ? str * (hq * g.red_xy) : str * (hq * m * g.red_xy);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            if (i > d.ied || j > d.jed) continue;
            const int li = tx + 1, lj = ty + 4 * r + 1;
            double du = 0.0, dv = 0.0;
            if (i >= is - 1 && i <= ie && j >= js && j <= je) {   // (:1059-1067)
                const double h_u = 0.5 * (s_mT[lj][li] * s_h[lj][li] + s_mT[lj][li + 1] * s_h[lj][li + 1]);
                du = ((g.IdxCu * ((g.dx2q * s_xy[lj - 1][li]) - (g.dx2q * s_xy[lj][li])) +
                       g.IdyCu * ((g.dy2h * s_xx[lj][li]) - (g.dy2h * s_xx[lj][li + 1]))) *
                      g.IareaCu) / (h_u + p.h_neglect);
            }
            if (i >= is && i <= ie && j >= js - 1 && j <= je) {   // (:1070-1078)
                const double h_v = 0.5 * (s_mT[lj][li] * s_h[lj][li] + s_mT[lj + 1][li] * s_h[lj + 1][li]);
                dv = ((g.IdyCv * ((g.dy2q * s_xy[lj][li - 1]) - (g.dy2q * s_xy[lj][li])) -
                       g.IdxCv * ((g.dx2h * s_xx[lj][li]) - (g.dx2h * s_xx[lj + 1][li]))) *
                      g.IareaCv) / (h_v + p.h_neglect);
            }
            const long long c = IX3(d, i, j, k);
            diffu[c] = du;
            diffv[c] = dv;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The general kernel: Smagorinsky and Leith viscosities and the
// thickness-aware bounds (:612-1040).  The reference fills 2D work arrays per
// layer; here every value is a function of the layer's u, v, h evaluated where
// it is needed, with the reference's ranges: a work value outside the range
// its loop writes reads as the array's initial value (zero; hrat_min and
// visc_bound_rem: one).  The Kh and Ah arrays are shared between h and q
// points: the h-point pass writes is..ie x js..je, the q-point pass is-1..ie x
// js-1..je, and the stress at the h points of the ring (is-1..ie+1 x
// js-1..je+1 minus the interior) reads what the previous layer's q-point pass
// left there -- for the first layer, the previous call's last layer (ring_Kh,
// ring_Ah), zero elsewhere.

__device__ inline double sq(double x) { return x * x; }

// inv_PI3 = 1/(3.14159265358979_dp**3), inv_PI6 = 1/(3.14159265358979_dp**6)
// (:154-155), the integer powers by squaring as the compiled reference has them
#define HV_PI 3.14159265358979
#define HV_INV_PI3 (1.0 / ((HV_PI * HV_PI) * HV_PI))
#define HV_INV_PI6 (1.0 / (((HV_PI * HV_PI) * (HV_PI * HV_PI)) * (HV_PI * HV_PI)))

__device__ inline bool interior(const Dims &d, int i, int j)
{
    return i >= d.isc && i <= d.iec && j >= d.jsc && j <= d.jec;
}
__device__ inline bool q_range(const Dims &d, int I, int J)   // Isq..Ieq x Jsq..Jeq
{
    return I >= d.isc - 1 && I <= d.iec && J >= d.jsc - 1 && J <= d.jec;
}

__device__ inline double dvdx_at(const HV &p, const double *v, int I, int J, int k)
{
    const Dims &d = p.d;
    return p.DY_dxBu[IX2(d, I, J)] *
           ((v[IX3(d, I + 1, J, k)] * p.IdyCv[IX2(d, I + 1, J)]) - (v[IX3(d, I, J, k)] * p.IdyCv[IX2(d, I, J)]));
}
__device__ inline double dudy_at(const HV &p, const double *u, int I, int J, int k)
{
    const Dims &d = p.d;
    return p.DX_dyBu[IX2(d, I, J)] *
           ((u[IX3(d, I, J + 1, k)] * p.IdxCu[IX2(d, I, J + 1)]) - (u[IX3(d, I, J, k)] * p.IdxCu[IX2(d, I, J)]));
}

// Leith: vorticity at q points (:612-626) and its gradients (:629-645)
__device__ inline double vort_at(const HV &p, const double *u, const double *v, int I, int J, int k)
{
    const double m = p.mask2dBu[IX2(p.d, I, J)];
    return (p.no_slip ? (2.0 - m) : m) * (dvdx_at(p, v, I, J, k) - dudy_at(p, u, I, J, k));
}
__device__ inline double vort_dx_at(const HV &p, const double *u, const double *v, int i, int j, int k)
{
    const Dims &d = p.d;
    if (!interior(d, i, j)) return 0.0;
    const double DY_dxBu_loc = p.dyBu[IX2(d, i, j)] * p.IdxBu[IX2(d, i, j)];
    return DY_dxBu_loc * ((vort_at(p, u, v, i, j, k) * p.IdyCu[IX2(d, i, j)]) -
                          (vort_at(p, u, v, i - 1, j, k) * p.IdyCu[IX2(d, i - 1, j)]));
}
__device__ inline double vort_dy_at(const HV &p, const double *u, const double *v, int I, int j, int k)
{
    const Dims &d = p.d;
    if (!(I >= d.isc - 1 && I <= d.iec && j >= d.jsc && j <= d.jec)) return 0.0;
    const double DX_dyBu_loc = p.dxBu[IX2(d, I, j)] * p.IdyBu[IX2(d, I, j)];
    return DX_dyBu_loc * ((vort_at(p, u, v, I, j, k) * p.IdxCv[IX2(d, I, j)]) -
                          (vort_at(p, u, v, I, j - 1, k) * p.IdxCv[IX2(d, I, j - 1)]));
}
// Laplacian of the vorticity at q points (:648-659), zero outside Isq..Ieq x Jsq..Jeq
__device__ inline double del2vort_q_at(const HV &p, const double *u, const double *v, int I, int J, int k)
{
    const Dims &d = p.d;
    if (!q_range(d, I, J)) return 0.0;
    return p.DY_dxBu[IX2(d, I, J)] * ((vort_dx_at(p, u, v, I + 1, J, k) * p.IdyCv[IX2(d, I + 1, J)]) -
                                      (vort_dx_at(p, u, v, I, J, k) * p.IdyCv[IX2(d, I, J)])) +
           p.DX_dyBu[IX2(d, I, J)] * ((vort_dy_at(p, u, v, I, J + 1, k) * p.IdxCu[IX2(d, I, J + 1)]) -
                                      (vort_dy_at(p, u, v, I, J, k) * p.IdxCu[IX2(d, I, J)]));
}

// thickness at u and v points with the land mask (:551-562)
__device__ inline double h_u_at(const HV &p, const double *h, int I, int j, int k)
{
    const Dims &d = p.d;
    return 0.5 * (p.mask2dT[IX2(d, I, j)] * h[IX3(d, I, j, k)] + p.mask2dT[IX2(d, I + 1, j)] * h[IX3(d, I + 1, j, k)]);
}
__device__ inline double h_v_at(const HV &p, const double *h, int i, int J, int k)
{
    const Dims &d = p.d;
    return 0.5 * (p.mask2dT[IX2(d, i, J)] * h[IX3(d, i, J, k)] + p.mask2dT[IX2(d, i, J + 1)] * h[IX3(d, i, J + 1, k)]);
}
// hrat_min (:730-737) on is..ie x js..je, one elsewhere
__device__ inline double hrat_at(const HV &p, const double *h, int i, int j, int k)
{
    if (!interior(p.d, i, j)) return 1.0;
    const double h_min = fmin(fmin(fmin(h_u_at(p, h, i, j, k), h_u_at(p, h, i - 1, j, k)), h_v_at(p, h, i, j, k)),
                              h_v_at(p, h, i, j - 1, k));
    return fmin(1.0, h_min / (h[IX3(p.d, i, j, k)] + p.h_neglect));
}

// Kh and Ah at an interior h point (:743-863); *rem = visc_bound_rem
__device__ inline void visc_h_point(const HV &p, const double *u, const double *v, const double *h, int i, int j,
                                    int k, double *Kh_out, double *Ah_out)
{
    const Dims &d = p.d;
    const double a2 = p.areaT[IX2(d, i, j)];
    double shear = 0.0, vvm = 0.0, hrat = 1.0;
    if (p.smag_Kh || p.smag_Ah) {   // Shear_mag (:710-720)
        const double sh_xx_sq = sq(sh_xx_at(p, u, v, i, j, k));
        const double sh_xy_sq = 0.25 * ((sq(sh_xy_at(p, u, v, i - 1, j - 1, k)) + sq(sh_xy_at(p, u, v, i, j, k))) +
                                        (sq(sh_xy_at(p, u, v, i - 1, j, k)) + sq(sh_xy_at(p, u, v, i, j - 1, k))));
        shear = sqrt(sh_xx_sq + sh_xy_sq);
    }
    if (p.leith_Kh) {   // vert_vort_mag = grad_vort_mag_h (+ 0: no modified Leith)
        const double gx = 0.5 * (vort_dx_at(p, u, v, i, j, k) + vort_dx_at(p, u, v, i, j - 1, k));
        const double gy = 0.5 * (vort_dy_at(p, u, v, i, j, k) + vort_dy_at(p, u, v, i - 1, j, k));
        vvm = sqrt(sq(gx) + sq(gy)) + 0.0;
    }
    if (p.bound) hrat = hrat_at(p, h, i, j, k);
    double rem = 1.0, Kh = 0.0;
    if (p.laplacian) {
        Kh = p.Kh;
        if (p.smag_Kh) Kh = fmax(Kh, 0.15 * 0.15 * a2 * shear);
        if (p.leith_Kh) Kh = fmax(Kh, 1.0 * (sqrt(a2) * a2) * vvm * HV_INV_PI3);
        Kh = fmax(Kh, 0.0);
        if (p.bound) {   // better_bound_Kh and better_bound_Ah together
            const double Kh_max_here = hrat * (0.25 * a2 / 1.0);
            if (Kh >= Kh_max_here) {
                rem = 0.0;
                Kh = Kh_max_here;
            } else if (Kh > 0.0) {
                rem = 1.0 - Kh / Kh_max_here;
            }
        }
    }
    *Kh_out = Kh;
    if (!p.biharmonic) return;
    double Ah = p.Ah;
    if (p.smag_Ah) Ah = fmax(Ah, 0.06 * 0.06 * (a2 * a2) * shear);
    if (p.leith_Ah) {
        const double Del2vort_h = 0.25 * ((del2vort_q_at(p, u, v, i, j, k) + del2vort_q_at(p, u, v, i - 1, j - 1, k)) +
                                          (del2vort_q_at(p, u, v, i - 1, j, k) + del2vort_q_at(p, u, v, i, j - 1, k)));
        const double h3 = sqrt(a2) * a2;
        Ah = fmax(Ah, 1.0 * h3 * h3 * fabs(Del2vort_h) * HV_INV_PI6);
    }
    if (p.bound) Ah = fmin(Ah, rem * hrat * (0.0625 * (a2 * a2) / 1.0));
    *Ah_out = Ah;
}

// Kh (which = 0) or Ah (1) at a q point of Isq..Ieq x Jsq..Jeq (:897-1023)
__device__ inline double visc_q_point(const HV &p, const double *u, const double *v, const double *h, int I, int J,
                                      int k, int which)
{
    const Dims &d = p.d;
    const double a2 = p.areaT[IX2(d, I, J)];
    const bool smag = which == 0 ? p.smag_Kh : p.smag_Ah;
    const bool leith = which == 0 ? p.leith_Kh : p.leith_Ah;
    double x = which == 0 ? p.Kh : p.Ah;
    if (smag) {
        const double sh_xx_sq = 0.25 * ((sq(sh_xx_at(p, u, v, I, J, k)) + sq(sh_xx_at(p, u, v, I + 1, J + 1, k))) +
                                        (sq(sh_xx_at(p, u, v, I + 1, J, k)) + sq(sh_xx_at(p, u, v, I, J + 1, k))));
        const double sh_xy_sq = sq(sh_xy_at(p, u, v, I, J, k));
        x = fmax(x, (which == 0 ? 0.15 * 0.15 * a2 : 0.06 * 0.06 * (a2 * a2)) * sqrt(sh_xx_sq + sh_xy_sq));
    }
    if (leith) {
        const double h3 = sqrt(a2) * a2;
        if (which == 0) {
            const double gx = 0.5 * (vort_dx_at(p, u, v, I, J, k) + vort_dx_at(p, u, v, I + 1, J, k));
            const double gy = 0.5 * (vort_dy_at(p, u, v, I, J, k) + vort_dy_at(p, u, v, I, J + 1, k));
            x = fmax(x, 1.0 * h3 * sqrt(sq(gx) + sq(gy)) * HV_INV_PI3);
        } else {
            x = fmax(x, 1.0 * h3 * h3 * fabs(del2vort_q_at(p, u, v, I, J, k)) * HV_INV_PI6);
        }
    }
    if (which == 0) x = fmax(x, 0.0);   // Kh_bg_min
    if (p.bound) {
        const double h_min = 0.25 * ((hrat_at(p, h, I, J, k) + hrat_at(p, h, I + 1, J + 1, k)) +
                                     (hrat_at(p, h, I + 1, J, k) + hrat_at(p, h, I, J + 1, k)));
        x = fmin(x, h_min * (which == 0 ? 0.25 * a2 / 1.0 : 0.0625 * (a2 * a2) / 1.0));
    }
    return x;
}

// Kh and Ah seen by the stress at h point (i, j) of Isq..Ieq+1 x Jsq..Jeq+1
__device__ inline void visc_seen(const HV &p, const double *u, const double *v, const double *h, int i, int j, int k,
                                 double *Kh, double *Ah)
{
    const Dims &d = p.d;
    if (interior(d, i, j)) {
        visc_h_point(p, u, v, h, i, j, k, Kh, Ah);
    } else if (q_range(d, i, j)) {
        if (k > 1) {
            *Kh = p.laplacian ? visc_q_point(p, u, v, h, i, j, k - 1, 0) : 0.0;
            *Ah = p.biharmonic ? visc_q_point(p, u, v, h, i, j, k - 1, 1) : 0.0;
        } else {
            *Kh = p.ring_Kh[IX2(d, i, j)];
            *Ah = p.ring_Ah[IX2(d, i, j)];
        }
    } else {
        *Kh = 0.0;
        *Ah = 0.0;
    }
}

__device__ inline double str_xx_x(const HV &p, const double *u, const double *v, const double *h, int i, int j, int k)
{
    const Dims &d = p.d;
    double Kh = 0.0, Ah = 0.0;
    visc_seen(p, u, v, h, i, j, k, &Kh, &Ah);
    double str = p.laplacian ? -(Kh * sh_xx_at(p, u, v, i, j, k)) : 0.0;
    if (p.biharmonic) {
        const double d_del2u = (p.IdyCu[IX2(d, i, j)] * p.Del2u[IX3(d, i, j, k)]) -
                               (p.IdyCu[IX2(d, i - 1, j)] * p.Del2u[IX3(d, i - 1, j, k)]);
        const double d_del2v = (p.IdxCv[IX2(d, i, j)] * p.Del2v[IX3(d, i, j, k)]) -
                               (p.IdxCv[IX2(d, i, j - 1)] * p.Del2v[IX3(d, i, j - 1, k)]);
        str = str + Ah * ((p.DY_dxT[IX2(d, i, j)] * d_del2u) - (p.DX_dyT[IX2(d, i, j)] * d_del2v));
    }
    return str * (h[IX3(d, i, j, k)] * p.reduction_xx[IX2(d, i, j)]);
}

__device__ inline double str_xy_x(const HV &p, const double *u, const double *v, const double *h, int I, int J, int k)
{
    const Dims &d = p.d;
    const long long c2 = IX2(d, I, J);
    double str = p.laplacian ? -(visc_q_point(p, u, v, h, I, J, k, 0) * sh_xy_at(p, u, v, I, J, k)) : 0.0;
    if (p.biharmonic) {
        const double d_str =
            visc_q_point(p, u, v, h, I, J, k, 1) *
            ((p.DY_dxBu[c2] * ((p.Del2v[IX3(d, I + 1, J, k)] * p.IdyCv[IX2(d, I + 1, J)]) -
                               (p.Del2v[IX3(d, I, J, k)] * p.IdyCv[c2]))) +
             (p.DX_dyBu[c2] * ((p.Del2u[IX3(d, I, J + 1, k)] * p.IdxCu[IX2(d, I, J + 1)]) -
                               (p.Del2u[IX3(d, I, J, k)] * p.IdxCu[c2]))));
        str = str + d_str;
    }
    const double hq = 0.25 * ((h[IX3(d, I, J, k)] + h[IX3(d, I + 1, J + 1, k)]) +
                              (h[IX3(d, I + 1, J, k)] + h[IX3(d, I, J + 1, k)]));
    if (p.no_slip) return str * (hq * p.reduction_xy[c2]);
    return str * (hq * p.mask2dBu[c2] * p.reduction_xy[c2]);
}

// diffu, diffv with the closures (the same ranges and expressions as k_hv_main)
__global__ void __launch_bounds__(64) k_hv_closures(HV p, const double *__restrict__ u, const double *__restrict__ v,
                                                    const double *__restrict__ h, double *__restrict__ diffu,
                                                    double *__restrict__ diffv)
{
    const Dims &d = p.d;
    const int i = d.isd + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsd + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.ied) return;
    const long long c = IX3(d, i, j, k), c2 = IX2(d, i, j);
    double du = 0.0, dv = 0.0;
    if (i >= d.isc - 1 && i <= d.iec && j >= d.jsc && j <= d.jec) {
        du = ((p.IdxCu[c2] * ((p.dx2q[IX2(d, i, j - 1)] * str_xy_x(p, u, v, h, i, j - 1, k)) -
                              (p.dx2q[c2] * str_xy_x(p, u, v, h, i, j, k))) +
               p.IdyCu[c2] * ((p.dy2h[c2] * str_xx_x(p, u, v, h, i, j, k)) -
                              (p.dy2h[IX2(d, i + 1, j)] * str_xx_x(p, u, v, h, i + 1, j, k)))) *
              p.IareaCu[c2]) / (h_u_at(p, h, i, j, k) + p.h_neglect);
    }
    if (i >= d.isc && i <= d.iec && j >= d.jsc - 1 && j <= d.jec) {
        dv = ((p.IdyCv[c2] * ((p.dy2q[IX2(d, i - 1, j)] * str_xy_x(p, u, v, h, i - 1, j, k)) -
                              (p.dy2q[c2] * str_xy_x(p, u, v, h, i, j, k))) -
               p.IdxCv[c2] * ((p.dx2h[c2] * str_xx_x(p, u, v, h, i, j, k)) -
                              (p.dx2h[IX2(d, i, j + 1)] * str_xx_x(p, u, v, h, i, j + 1, k)))) *
              p.IareaCv[c2]) / (h_v_at(p, h, i, j, k) + p.h_neglect);
    }
    diffu[c] = du;
    diffv[c] = dv;
}

// After a call: the q-point Kh and Ah of its last layer at the ring points
// (is-1, js-1..je) and (is..ie, js-1), for the next call's first layer
__global__ void k_hv_keep_ring(HV p, const double *__restrict__ u, const double *__restrict__ v,
                               const double *__restrict__ h)
{
    const Dims &d = p.d;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = d.jec - d.jsc + 2, nrow = d.iec - d.isc + 1;
    if (q >= ncol + nrow) return;
    const int I = q < ncol ? d.isc - 1 : d.isc + (q - ncol);
    const int J = q < ncol ? d.jsc - 1 + q : d.jsc - 1;
    p.ring_Kh[IX2(d, I, J)] = p.laplacian ? visc_q_point(p, u, v, h, I, J, d.nk, 0) : 0.0;
    p.ring_Ah[IX2(d, I, J)] = p.biharmonic ? visc_q_point(p, u, v, h, I, J, d.nk, 1) : 0.0;
}

static HV make_hv(const mom6h_hor_visc_cs *cs)
{
    HV p;
    p.d = cs->d;
    p.Kh = cs->Kh; p.Ah = cs->Ah; p.h_neglect = cs->h_neglect;
    p.laplacian = cs->laplacian; p.biharmonic = cs->biharmonic; p.no_slip = cs->no_slip;
    p.first_call = cs->calls == 0;
    p.DY_dxT = cs->DY_dxT; p.DX_dyT = cs->DX_dyT; p.DY_dxBu = cs->DY_dxBu; p.DX_dyBu = cs->DX_dyBu;
    p.IdyCu = cs->IdyCu; p.IdxCu = cs->IdxCu; p.IdyCv = cs->IdyCv; p.IdxCv = cs->IdxCv;
    p.IareaCu = cs->IareaCu; p.IareaCv = cs->IareaCv; p.mask2dT = cs->mask2dT; p.mask2dBu = cs->mask2dBu;
    p.reduction_xx = cs->reduction_xx; p.reduction_xy = cs->reduction_xy;
    p.dy2h = cs->dy2h; p.dx2h = cs->dx2h; p.dy2q = cs->dy2q; p.dx2q = cs->dx2q;
    p.Idx2dyCu = cs->Idx2dyCu; p.Idxdy2u = cs->Idxdy2u; p.Idx2dyCv = cs->Idx2dyCv; p.Idxdy2v = cs->Idxdy2v;
    p.Del2u = cs->Del2u; p.Del2v = cs->Del2v;
    p.smag_Kh = cs->smagorinsky && cs->laplacian;
    p.smag_Ah = cs->smagorinsky && cs->biharmonic;
    p.leith_Kh = cs->leith;
    p.leith_Ah = cs->leith && cs->biharmonic;
    p.bound = cs->better_bound;
    p.areaT = cs->areaT; p.dxBu = cs->dxBu; p.dyBu = cs->dyBu; p.IdxBu = cs->IdxBu; p.IdyBu = cs->IdyBu;
    p.ring_Kh = cs->ring_Kh; p.ring_Ah = cs->ring_Ah;
    return p;
}

static bool closures_on(const mom6h_hor_visc_cs *cs) { return cs->smagorinsky || cs->leith || cs->better_bound; }

// the tiled kernel's launch: 8-row tiles over the data domain, the layers
// split so that a block walks about 15 of them
static void launch_tile(const mom6h_hor_visc_cs *cs, const HV &p, const double *u, const double *v,
                        const double *h, double *diffu, double *diffv, hipStream_t s)
{
    const Dims &d = cs->d;
    constexpr int TY = 8;
    const int nkz = (d.nk + 14) / 15;
    const dim3 grid((unsigned)((d.n1 + HT_X - 1) / HT_X), (unsigned)((d.n2 + TY - 1) / TY), (unsigned)nkz);
    if (cs->biharmonic) k_hv_tile<true, TY><<<grid, 256, 0, s>>>(p, cs->g, u, v, h, diffu, diffv);
    else k_hv_tile<false, TY><<<grid, 256, 0, s>>>(p, cs->g, u, v, h, diffu, diffv);
}

// every metric the stencil reads constant over is-2..ie+2 x js-2..je+2
static bool hv_uniform(const Dims &d, std::initializer_list<std::pair<const double *, double *>> fields)
{
    for (const auto &f : fields)
        if (!uniform_over(d, f.first, d.isc - 2, d.iec + 2, d.jsc - 2, d.jec + 2, f.second)) return false;
    return true;
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_hor_visc_init(mom6h_hor_visc_cs **out, int isd, int ied, int jsd, int jed, int isc,
                        int iec, int jsc, int jec, int nz, double Kh, double h_neglect,
                        const double *DY_dxT, const double *DX_dyT, const double *DY_dxBu,
                        const double *DX_dyBu, const double *IdyCu, const double *IdxCu,
                        const double *IdyCv, const double *IdxCv, const double *IareaCu,
                        const double *IareaCv, const double *mask2dT, const double *mask2dBu,
                        const double *reduction_xx, const double *reduction_xy, const double *dy2h,
                        const double *dx2h, const double *dy2q, const double *dx2q)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_hor_visc_init: NULL cs");
    MOM6H_REQUIRE(isd < isc - 1 && iec + 1 < ied && jsd < jsc - 1 && jec + 1 < jed && isc <= iec &&
                      jsc <= jec && nz >= 1,
                  "mom6h_hor_visc_init: needs two halo cells on every side");
    MOM6H_REQUIRE(DY_dxT && DX_dyT && DY_dxBu && DX_dyBu && IdyCu && IdxCu && IdyCv && IdxCv &&
                      IareaCu && IareaCv && mask2dT && mask2dBu && reduction_xx && reduction_xy &&
                      dy2h && dx2h && dy2q && dx2q,
                  "mom6h_hor_visc_init: NULL metric");
    auto *cs = new mom6h_hor_visc_cs();
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nz);
    cs->Kh = Kh;
    cs->Ah = 0.0;
    cs->h_neglect = h_neglect;
    cs->laplacian = 1;
    cs->biharmonic = 0;
    cs->no_slip = 0;
    cs->calls = 0;
    const size_t n2 = (size_t)cs->d.s2;
    int rc = 0;
    const double *src[] = {DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, IdxCv, IareaCu,
                           IareaCv, mask2dT, mask2dBu, reduction_xx, reduction_xy, dy2h, dx2h, dy2q, dx2q};
    double **dst[] = {&cs->DY_dxT, &cs->DX_dyT, &cs->DY_dxBu, &cs->DX_dyBu, &cs->IdyCu, &cs->IdxCu,
                      &cs->IdyCv, &cs->IdxCv, &cs->IareaCu, &cs->IareaCv, &cs->mask2dT, &cs->mask2dBu,
                      &cs->reduction_xx, &cs->reduction_xy, &cs->dy2h, &cs->dx2h, &cs->dy2q, &cs->dx2q};
    for (size_t q = 0; q < sizeof(src) / sizeof(src[0]); ++q) rc = rc ? rc : dev_upload(dst[q], src[q], n2);
    if (rc) {
        mom6h_hor_visc_end(cs);
        return rc;
    }
    HVUniform &g = cs->g;
    cs->uniform = hv_uniform(cs->d, {{DY_dxT, &g.DY_dxT}, {DX_dyT, &g.DX_dyT}, {DY_dxBu, &g.DY_dxBu},
                                     {DX_dyBu, &g.DX_dyBu}, {IdyCu, &g.IdyCu}, {IdxCu, &g.IdxCu},
                                     {IdyCv, &g.IdyCv}, {IdxCv, &g.IdxCv}, {IareaCu, &g.IareaCu},
                                     {IareaCv, &g.IareaCv}, {reduction_xx, &g.red_xx},
                                     {reduction_xy, &g.red_xy}, {dy2h, &g.dy2h}, {dx2h, &g.dx2h},
                                     {dy2q, &g.dy2q}, {dx2q, &g.dx2q}});
    *out = cs;
    return MOM6H_OK;
}

int mom6h_hor_visc_set_options(mom6h_hor_visc_cs *cs, int laplacian, int biharmonic, double Ah,
                               int no_slip, const double *Idx2dyCu, const double *Idxdy2u,
                               const double *Idx2dyCv, const double *Idxdy2v)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_hor_visc_set_options: NULL cs");
    MOM6H_REQUIRE(!biharmonic || (Idx2dyCu && Idxdy2u && Idx2dyCv && Idxdy2v),
                  "mom6h_hor_visc_set_options: biharmonic needs its four metrics");
    cs->laplacian = laplacian;
    cs->no_slip = no_slip;
    cs->Ah = Ah;
    if (biharmonic && !cs->biharmonic) {
        const size_t n2 = (size_t)cs->d.s2, n3 = n2 * cs->d.nk;
        int rc = 0;
        rc = rc ? rc : dev_upload(&cs->Idx2dyCu, Idx2dyCu, n2);
        rc = rc ? rc : dev_upload(&cs->Idxdy2u, Idxdy2u, n2);
        rc = rc ? rc : dev_upload(&cs->Idx2dyCv, Idx2dyCv, n2);
        rc = rc ? rc : dev_upload(&cs->Idxdy2v, Idxdy2v, n2);
        rc = rc ? rc : dev_alloc(&cs->Del2u, n3, true);
        rc = rc ? rc : dev_alloc(&cs->Del2v, n3, true);
        if (rc) return rc;
        HVUniform &g = cs->g;
        cs->bih_uniform = hv_uniform(cs->d, {{Idx2dyCu, &g.Idx2dyCu}, {Idxdy2u, &g.Idxdy2u},
                                             {Idx2dyCv, &g.Idx2dyCv}, {Idxdy2v, &g.Idxdy2v}});
    }
    cs->biharmonic = biharmonic;
    return MOM6H_OK;
}

int mom6h_hor_visc(mom6h_hor_visc_cs *cs, const double *u, const double *v, const double *h,
                   double *diffu, double *diffv, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_hor_visc: NULL cs");
    MOM6H_REQUIRE(u && v && h && diffu && diffv, "mom6h_hor_visc: NULL array");
    const Dims &d = cs->d;
    hipStream_t s = (hipStream_t)stream;
    HV p = make_hv(cs);
    if (cs->biharmonic) {
        const int ni = d.iec - (d.isc - 1) + 1, nj = d.jec - (d.jsc - 1) + 1;
        dim3 b(64), g((ni + 63) / 64, nj, d.nk);
        k_hv_del2<<<g, b, 0, s>>>(p, u, v);
        MOM6H_LAUNCH_CHECK();
    }
    dim3 b(64), g((d.n1 + 63) / 64, d.n2, d.nk);
    if (closures_on(cs)) {
        k_hv_closures<<<g, b, 0, s>>>(p, u, v, h, diffu, diffv);
        MOM6H_LAUNCH_CHECK();
        const int nring = (d.jec - d.jsc + 2) + (d.iec - d.isc + 1);
        k_hv_keep_ring<<<(nring + 63) / 64, 64, 0, s>>>(p, u, v, h);
    } else if (cs->uniform && (!cs->biharmonic || cs->bih_uniform) && !cs->general) {
        launch_tile(cs, p, u, v, h, diffu, diffv, s);
    } else {
        k_hv_main<<<g, b, 0, s>>>(p, u, v, h, diffu, diffv);
    }
    MOM6H_LAUNCH_CHECK();
    cs->calls++;
    return MOM6H_OK;
}

int mom6h_hor_visc_set_closures(mom6h_hor_visc_cs *cs, int smagorinsky, int leith, int better_bound,
                                const double *areaT, const double *dxBu, const double *dyBu, const double *IdxBu,
                                const double *IdyBu)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_hor_visc_set_closures: NULL cs");
    MOM6H_REQUIRE(cs->calls == 0, "mom6h_hor_visc_set_closures: before the first mom6h_hor_visc call");
    MOM6H_REQUIRE(!(smagorinsky || leith || better_bound) || areaT, "mom6h_hor_visc_set_closures: needs areaT");
    MOM6H_REQUIRE(!leith || (dxBu && dyBu && IdxBu && IdyBu), "mom6h_hor_visc_set_closures: Leith needs the q-point metrics");
    const size_t n2 = (size_t)cs->d.s2;
    int rc = 0;
    if (smagorinsky || leith || better_bound) {
        if (!cs->areaT) rc = dev_upload(&cs->areaT, areaT, n2);
        if (!rc && !cs->ring_Kh) rc = dev_alloc(&cs->ring_Kh, n2, true);
        if (!rc && !cs->ring_Ah) rc = dev_alloc(&cs->ring_Ah, n2, true);
    }
    if (!rc && leith && !cs->dxBu) {
        rc = dev_upload(&cs->dxBu, dxBu, n2);
        rc = rc ? rc : dev_upload(&cs->dyBu, dyBu, n2);
        rc = rc ? rc : dev_upload(&cs->IdxBu, IdxBu, n2);
        rc = rc ? rc : dev_upload(&cs->IdyBu, IdyBu, n2);
    }
    if (rc) return rc;
    cs->smagorinsky = smagorinsky;
    cs->leith = leith;
    cs->better_bound = better_bound;
    return MOM6H_OK;
}

int mom6h_hor_visc_end(mom6h_hor_visc_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->DY_dxT, cs->DX_dyT, cs->DY_dxBu, cs->DX_dyBu, cs->IdyCu, cs->IdxCu,
                      cs->IdyCv, cs->IdxCv, cs->IareaCu, cs->IareaCv, cs->mask2dT, cs->mask2dBu,
                      cs->reduction_xx, cs->reduction_xy, cs->dy2h, cs->dx2h, cs->dy2q, cs->dx2q,
                      cs->Idx2dyCu, cs->Idxdy2u, cs->Idx2dyCv, cs->Idxdy2v, cs->Del2u, cs->Del2v,
                      cs->areaT, cs->dxBu, cs->dyBu, cs->IdxBu, cs->IdyBu, cs->ring_Kh, cs->ring_Ah})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
