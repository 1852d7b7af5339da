// mom6h_hor_visc.hip -- horizontal viscosity on gfx950.
//
// Arithmetic of src/serial/mom6_hor_visc_serial.F90:464-1109 with the
// Laplacian (constant Kh) and biharmonic (constant Ah) operators, free- or
// no-slip.  One 3D launch computes every layer: each thread forms the
// strain/stress values its diffu/diffv stencil needs directly from u, v, h
// (no 2D work arrays, no per-layer launches).
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
    return p;
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
    k_hv_main<<<g, b, 0, s>>>(p, u, v, h, diffu, diffv);
    MOM6H_LAUNCH_CHECK();
    cs->calls++;
    return MOM6H_OK;
}

int mom6h_hor_visc_end(mom6h_hor_visc_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->DY_dxT, cs->DX_dyT, cs->DY_dxBu, cs->DX_dyBu, cs->IdyCu, cs->IdxCu,
                      cs->IdyCv, cs->IdxCv, cs->IareaCu, cs->IareaCv, cs->mask2dT, cs->mask2dBu,
                      cs->reduction_xx, cs->reduction_xy, cs->dy2h, cs->dx2h, cs->dy2q, cs->dx2q,
                      cs->Idx2dyCu, cs->Idxdy2u, cs->Idx2dyCv, cs->Idxdy2v, cs->Del2u, cs->Del2v})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
