// mom6h_coriolis.hip -- CorAdCalc on gfx950.
//
// Arithmetic of src/serial/mom6_coriolis_serial.F90:112-252.  Two stages per
// call, each a 3D launch (all layers at once):
//   k_cor_qke : q on I in is-1..ie, J in js-1..je and KE on is..ie x js..je
//   k_cor_ca  : CAu on is..ie-1 x js..je, CAv on is..ie x js..je-1
// The serial code never writes hArea_u on row js-1 nor hArea_v on column
// is-1 (computed ranges :142-152) and reads them as zero when forming q on the
// southern row / western column (:162-168); the same zeros are used here.
#include "mom6h_internal.h"

namespace mom6h {

__global__ void k_cor_qke(Dims d, const double *__restrict__ u, const double *__restrict__ v,
                          const double *__restrict__ h, const double *__restrict__ areaT,
                          const double *__restrict__ IareaBu, const double *__restrict__ CoriolisBu,
                          const double *__restrict__ mask2dBu, const double *__restrict__ dyCv,
                          const double *__restrict__ dxCu, const double *__restrict__ dyCu,
                          const double *__restrict__ dxCv, const double *__restrict__ Area_q,
                          double *__restrict__ q, double *__restrict__ KE)
{
    const int i = d.isc - 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc - 1 + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.iec) return;
    const int is = d.isc, ie = d.iec, js = d.jsc;
    const long long c = IX3(d, i, j, k);
    // circulation (:134-139)
    const double dvdx = (v[IX3(d, i + 1, j, k)] * dyCv[IX2(d, i + 1, j)]) - (v[c] * dyCv[IX2(d, i, j)]);
    const double dudy = (u[IX3(d, i, j + 1, k)] * dxCu[IX2(d, i, j + 1)]) - (u[c] * dxCu[IX2(d, i, j)]);
    // thickness-weighted areas (:142-152); out-of-range entries are the zeros the reference reads
    auto hArea_u = [&](int p, int r) -> double {
        if (r < js || r > d.jec + 1 || p < is - 1 || p > ie) return 0.0;
        return 0.5 * ((areaT[IX2(d, p, r)] * h[IX3(d, p, r, k)]) + (areaT[IX2(d, p + 1, r)] * h[IX3(d, p + 1, r, k)]));
    };
    auto hArea_v = [&](int p, int r) -> double {
        if (p < is || p > ie + 1 || r < js - 1 || r > d.jec) return 0.0;
        return 0.5 * ((areaT[IX2(d, p, r)] * h[IX3(d, p, r, k)]) + (areaT[IX2(d, p, r + 1)] * h[IX3(d, p, r + 1, k)]));
    };
    const long long c2 = IX2(d, i, j);
    const double rel_vort = mask2dBu[c2] * (dvdx - dudy) * IareaBu[c2];
    const double abs_vort = CoriolisBu[c2] + rel_vort;
    const double hArea_q = (hArea_u(i, j) + hArea_u(i, j + 1)) + (hArea_v(i, j) + hArea_v(i + 1, j));
    const double Ih_q = Area_q[c2] / (hArea_q + 1.0e-20);
    q[c] = abs_vort * Ih_q;
    // kinetic energy (:196-203)
    if (i >= is && j >= js) {
        const double uc = u[c], uw = u[IX3(d, i - 1, j, k)];
        const double vc = v[c], vs = v[IX3(d, i, j - 1, k)];
        KE[c] = 0.25 * ((dyCu[c2] * (uc * uc) + dyCu[IX2(d, i - 1, j)] * (uw * uw)) +
                        (dxCv[c2] * (vc * vc) + dxCv[IX2(d, i, j - 1)] * (vs * vs))) /
                areaT[c2];
    }
}

__global__ void k_cor_ca(Dims d, int scheme, const double *__restrict__ q, const double *__restrict__ KE,
                         const double *__restrict__ uh, const double *__restrict__ vh,
                         const double *__restrict__ IdxCu, const double *__restrict__ IdyCv,
                         double *__restrict__ CAu, double *__restrict__ CAv)
{
    const int i = d.isc + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.jsc + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > d.iec) return;
    const long long c = IX3(d, i, j, k);
    auto Q = [&](int p, int r) { return q[IX3(d, p, r, k)]; };
    const double C1_12 = 1.0 / 12.0, C1_24 = 1.0 / 24.0;
    const bool do_u = i <= d.iec - 1, do_v = j <= d.jec - 1;
    if (scheme == 1) {  // Sadourny (1975) energy conserving (:206-226)
        if (do_u) {
            const double KEx = (KE[IX3(d, i + 1, j, k)] - KE[c]) * IdxCu[IX2(d, i, j)];
            CAu[c] = 0.25 * ((Q(i, j) * (vh[IX3(d, i + 1, j, k)] + vh[c])) +
                             (Q(i, j - 1) * (vh[IX3(d, i, j - 1, k)] + vh[IX3(d, i + 1, j - 1, k)]))) *
                         IdxCu[IX2(d, i, j)] - KEx;
        }
        if (do_v) {
            const double KEy = (KE[IX3(d, i, j + 1, k)] - KE[c]) * IdyCv[IX2(d, i, j)];
            CAv[c] = -(0.25 * ((Q(i - 1, j) * (uh[IX3(d, i - 1, j, k)] + uh[IX3(d, i - 1, j + 1, k)])) +
                               (Q(i, j) * (uh[c] + uh[IX3(d, i, j + 1, k)]))) *
                       IdyCv[IX2(d, i, j)]) - KEy;
        }
        return;
    }
    // Arakawa & Hsu (1990) / Arakawa & Lamb (1981) coefficients (:171-193)
    auto A = [&](int p, int r) {
        return scheme == 2 ? (Q(p, r) + (Q(p + 1, r) + Q(p, r - 1))) * C1_12
                           : (2.0 * (Q(p + 1, r) + Q(p, r - 1)) + (Q(p, r) + Q(p + 1, r - 1))) * C1_24;
    };
    auto D = [&](int p, int r) {
        return scheme == 2 ? ((Q(p, r) + Q(p + 1, r - 1)) + Q(p, r - 1)) * C1_12
                           : ((Q(p + 1, r) + Q(p, r - 1)) + 2.0 * (Q(p, r) + Q(p + 1, r - 1))) * C1_24;
    };
    auto B = [&](int p, int r) {
        return scheme == 2 ? (Q(p, r) + (Q(p - 1, r) + Q(p, r - 1))) * C1_12
                           : ((Q(p, r) + Q(p - 1, r - 1)) + 2.0 * (Q(p - 1, r) + Q(p, r - 1))) * C1_24;
    };
    auto Cc = [&](int p, int r) {
        return scheme == 2 ? ((Q(p, r) + Q(p - 1, r - 1)) + Q(p, r - 1)) * C1_12
                           : (2.0 * (Q(p, r) + Q(p - 1, r - 1)) + (Q(p - 1, r) + Q(p, r - 1))) * C1_24;
    };
    if (do_u) {  // (:229-237)
        const double KEx = (KE[IX3(d, i + 1, j, k)] - KE[c]) * IdxCu[IX2(d, i, j)];
        CAu[c] = (((A(i, j) * vh[IX3(d, i + 1, j, k)]) + (Cc(i, j) * vh[IX3(d, i, j - 1, k)])) +
                  ((B(i, j) * vh[c]) + (D(i, j) * vh[IX3(d, i + 1, j - 1, k)]))) *
                     IdxCu[IX2(d, i, j)] - KEx;
    }
    if (do_v) {  // (:239-247)
        const double KEy = (KE[IX3(d, i, j + 1, k)] - KE[c]) * IdyCv[IX2(d, i, j)];
        CAv[c] = -((((A(i - 1, j) * uh[IX3(d, i - 1, j, k)]) + (Cc(i, j + 1) * uh[IX3(d, i, j + 1, k)])) +
                    ((B(i, j) * uh[c]) + (D(i - 1, j + 1) * uh[IX3(d, i - 1, j + 1, k)]))) *
                   IdyCv[IX2(d, i, j)]) - KEy;
    }
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_coriolis_init(mom6h_coriolis_cs **out, int isd, int ied, int jsd, int jed, int isc,
                        int iec, int jsc, int jec, int nk, const double *areaT,
                        const double *IareaBu, const double *CoriolisBu, const double *mask2dBu,
                        const double *dyCv, const double *dxCu, const double *dyCu,
                        const double *dxCv, const double *IdxCu, const double *IdyCv, int scheme)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_coriolis_init: NULL cs");
    MOM6H_REQUIRE(isd < isc && iec < ied && jsd < jsc && jec < jed && isc <= iec && jsc <= jec && nk >= 1,
                  "mom6h_coriolis_init: needs one halo cell on every side");
    MOM6H_REQUIRE(scheme >= 1 && scheme <= 3, "mom6h_coriolis_init: scheme must be 1, 2 or 3");
    MOM6H_REQUIRE(areaT && IareaBu && CoriolisBu && mask2dBu && dyCv && dxCu && dyCu && dxCv && IdxCu && IdyCv,
                  "mom6h_coriolis_init: NULL metric");
    auto *cs = new mom6h_coriolis_cs();
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, nk);
    cs->scheme = scheme;
    const size_t n2 = (size_t)cs->d.s2, n3 = n2 * nk;
    // coriolis_init Area_q (mom6_coriolis_serial.F90:74-80)
    double *aq = (double *)calloc(n2, sizeof(double));
    const Dims &d = cs->d;
    for (int j = jsd; j <= jed - 1; ++j)
        for (int i = isd; i <= ied - 1; ++i)
            aq[IX2(d, i, j)] = (areaT[IX2(d, i, j)] + areaT[IX2(d, i + 1, j + 1)]) +
                               (areaT[IX2(d, i + 1, j)] + areaT[IX2(d, i, j + 1)]);
    int rc = 0;
    rc = rc ? rc : dev_upload(&cs->areaT, areaT, n2);
    rc = rc ? rc : dev_upload(&cs->IareaBu, IareaBu, n2);
    rc = rc ? rc : dev_upload(&cs->CoriolisBu, CoriolisBu, n2);
    rc = rc ? rc : dev_upload(&cs->mask2dBu, mask2dBu, n2);
    rc = rc ? rc : dev_upload(&cs->dyCv, dyCv, n2);
    rc = rc ? rc : dev_upload(&cs->dxCu, dxCu, n2);
    rc = rc ? rc : dev_upload(&cs->dyCu, dyCu, n2);
    rc = rc ? rc : dev_upload(&cs->dxCv, dxCv, n2);
    rc = rc ? rc : dev_upload(&cs->IdxCu, IdxCu, n2);
    rc = rc ? rc : dev_upload(&cs->IdyCv, IdyCv, n2);
    rc = rc ? rc : dev_upload(&cs->Area_q, aq, n2);
    rc = rc ? rc : dev_alloc(&cs->q, n3, true);
    rc = rc ? rc : dev_alloc(&cs->KE, n3, true);
    free(aq);
    if (rc) {
        mom6h_coriolis_end(cs);
        return rc;
    }
    *out = cs;
    return MOM6H_OK;
}

int mom6h_coradcalc(mom6h_coriolis_cs *cs, const double *u, const double *v, const double *h,
                    const double *uh, const double *vh, double *CAu, double *CAv, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_coradcalc: NULL cs");
    MOM6H_REQUIRE(u && v && h && uh && vh && CAu && CAv, "mom6h_coradcalc: NULL array");
    const Dims &d = cs->d;
    hipStream_t s = (hipStream_t)stream;
    {
        const int ni = d.iec - (d.isc - 1) + 1, nj = d.jec - (d.jsc - 1) + 1;
        dim3 b(64), g((ni + 63) / 64, nj, d.nk);
        k_cor_qke<<<g, b, 0, s>>>(d, u, v, h, cs->areaT, cs->IareaBu, cs->CoriolisBu, cs->mask2dBu,
                                  cs->dyCv, cs->dxCu, cs->dyCu, cs->dxCv, cs->Area_q, cs->q, cs->KE);
        MOM6H_LAUNCH_CHECK();
    }
    {
        const int ni = d.iec - d.isc + 1, nj = d.jec - d.jsc + 1;
        dim3 b(64), g((ni + 63) / 64, nj, d.nk);
        k_cor_ca<<<g, b, 0, s>>>(d, cs->scheme, cs->q, cs->KE, uh, vh, cs->IdxCu, cs->IdyCv, CAu, CAv);
        MOM6H_LAUNCH_CHECK();
    }
    return MOM6H_OK;
}

int mom6h_coriolis_end(mom6h_coriolis_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->areaT, cs->IareaBu, cs->CoriolisBu, cs->mask2dBu, cs->dyCv, cs->dxCu,
                      cs->dyCu, cs->dxCv, cs->IdxCu, cs->IdyCv, cs->Area_q, cs->q, cs->KE})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
