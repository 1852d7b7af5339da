// mom6h_coriolis.hip -- CorAdCalc on gfx950.
//
// Arithmetic of src/serial/mom6_coriolis_serial.F90:112-252.  Two stages per
// call, each a 3D launch (all layers at once):
//   k_cor_qke : q on I in is-1..ie, J in js-1..je and KE on is..ie x js..je
//   k_cor_ca  : CAu on is..ie-1 x js..je, CAv on is..ie x js..je-1
// The serial code never writes hArea_u on row js-1 nor hArea_v on column
// is-1 (computed ranges :142-152) and reads them as zero when forming q on the
// southern row / western column (:162-168); the same zeros are used here.
// On grids whose metric arrays are constant both stages run as one tiled
// layer walk instead (k_cor_tile), bit-identical to the two general kernels.
#include "mom6h_internal.h"

#include <utility>

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

// ---------------------------------------------------------------------------
// Grids whose metric arrays are constant (over is-2..ie+2 x js-2..je+2, every
// point the stencil reads; Area_q over its q points): both stages in one
// tiled layer walk.  A block owns CT_X x TY points of the compute domain and
// walks a range of layers; per layer it stages u, v, h, uh, vh of the tile
// with a ring of one cell below and two above in LDS, forms q on the q points
// and KE on the h points the tile's CAu/CAv read, then CAu and CAv.  q and KE
// never leave the chip.  The expressions are k_cor_qke's and k_cor_ca's with
// the constant metric values, so CAu and CAv are bit-identical to theirs.
struct CorUniform {
    double areaT, IareaBu, dyCv, dxCu, dyCu, dxCv, IdxCu, IdyCv, Area_q;
};

constexpr int CT_X = 64, CW = CT_X + 3;

template <int SCHEME, int TY>
__global__ void __launch_bounds__(256) k_cor_tile(Dims d, CorUniform g, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ h,
                                                  const double *__restrict__ uh, const double *__restrict__ vh,
                                                  const double *__restrict__ CoriolisBu,
                                                  const double *__restrict__ mask2dBu, double *__restrict__ CAu,
                                                  double *__restrict__ CAv)
{
    constexpr int CH = TY + 3, RPT = TY / 4, NS = (CW * CH + 255) / 256;
    __shared__ double s_u[CH][CW], s_v[CH][CW], s_h[CH][CW], s_uh[CH][CW], s_vh[CH][CW];
    __shared__ double s_q[CH][CW], s_KE[CH][CW], s_cor[CH][CW], s_mBu[CH][CW];
    const int nxy = gridDim.x * gridDim.y;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nxy * gridDim.z);
    const int xy = tile % nxy, kz = tile / nxy;
    const int per = (d.nk + (int)gridDim.z - 1) / (int)gridDim.z;
    const int k0 = 1 + kz * per, k1 = k0 + per - 1 < d.nk ? k0 + per - 1 : d.nk;
    if (k0 > k1) return;
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const int i0 = is + (xy % (int)gridDim.x) * CT_X, j0 = js + (xy / (int)gridDim.x) * TY;
    const int bi = i0 - 1, bj = j0 - 1;   // staged region [i0-1, i1+2] x [j0-1, j1+2]
    const int tid = threadIdx.x, tx = tid % CT_X, ty = tid / CT_X;

    double pu[NS], pv[NS], ph[NS], puh[NS], pvh[NS];
    auto load_layer = [&](int k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r, li = t % CW, lj = t / CW, i = bi + li, j = bj + lj;
            pu[r] = 0.0; pv[r] = 0.0; ph[r] = 0.0; puh[r] = 0.0; pvh[r] = 0.0;
            if (t < CW * CH && i <= d.ied && j <= d.jed) {
                const long long c = IX3(d, i, j, k);
                pu[r] = u[c]; pv[r] = v[c]; ph[r] = h[c]; puh[r] = uh[c]; pvh[r] = vh[c];
            }
        }
    };
    for (int t = tid; t < CW * CH; t += 256) {
        const int li = t % CW, lj = t / CW, i = bi + li, j = bj + lj;
        double cor = 0.0, mb = 0.0;
        if (i <= d.ied && j <= d.jed) {
            cor = CoriolisBu[IX2(d, i, j)];
            mb = mask2dBu[IX2(d, i, j)];
        }
        s_cor[lj][li] = cor;
        s_mBu[lj][li] = mb;
    }
    load_layer(k0);
    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            if (t < CW * CH) {
                const int li = t % CW, lj = t / CW;
                s_u[lj][li] = pu[r]; s_v[lj][li] = pv[r]; s_h[lj][li] = ph[r];
                s_uh[lj][li] = puh[r]; s_vh[lj][li] = pvh[r];
            }
        }
        __syncthreads();
        if (k < k1) load_layer(k + 1);
        for (int t = tid; t < CW * CH; t += 256) {
            const int li = t % CW, lj = t / CW, I = bi + li, J = bj + lj;
            // q on I in is-1..ie, J in js-1..je (:134-168); hArea_u on row
            // js-1 and hArea_v on column is-1 are the never-written zeros
            if (li < CW - 1 && lj < CH - 1 && I <= ie && J <= je) {
                const double dvdx = (s_v[lj][li + 1] * g.dyCv) - (s_v[lj][li] * g.dyCv);
                const double dudy = (s_u[lj + 1][li] * g.dxCu) - (s_u[lj][li] * g.dxCu);
                const double hAu0 = J < js ? 0.0 : 0.5 * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj][li + 1]));
                const double hAu1 = 0.5 * ((g.areaT * s_h[lj + 1][li]) + (g.areaT * s_h[lj + 1][li + 1]));
                const double hAv0 = I < is ? 0.0 : 0.5 * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj + 1][li]));
                const double hAv1 = 0.5 * ((g.areaT * s_h[lj][li + 1]) + (g.areaT * s_h[lj + 1][li + 1]));
                const double rel_vort = s_mBu[lj][li] * (dvdx - dudy) * g.IareaBu;
                const double abs_vort = s_cor[lj][li] + rel_vort;
                const double hArea_q = (hAu0 + hAu1) + (hAv0 + hAv1);
                const double Ih_q = g.Area_q / (hArea_q + 1.0e-20);
                s_q[lj][li] = abs_vort * Ih_q;
            }
            // KE on is..ie x js..je (:196-203)
            if (li >= 1 && lj >= 1 && I >= is && I <= ie && J >= js && J <= je) {
                const double uc = s_u[lj][li], uw = s_u[lj][li - 1];
                const double vc = s_v[lj][li], vs = s_v[lj - 1][li];
                s_KE[lj][li] = 0.25 * ((g.dyCu * (uc * uc) + g.dyCu * (uw * uw)) +
                                       (g.dxCv * (vc * vc) + g.dxCv * (vs * vs))) /
                               g.areaT;
            }
        }
        __syncthreads();
        const double C1_12 = 1.0 / 12.0, C1_24 = 1.0 / 24.0;
        auto Q = [&](int pi, int pj) { return s_q[pj][pi]; };
        // Arakawa & Hsu (1990) / Arakawa & Lamb (1981) coefficients (:171-193)
        auto A = [&](int pi, int pj) {
            return SCHEME == 2 ? (Q(pi, pj) + (Q(pi + 1, pj) + Q(pi, pj - 1))) * C1_12
                               : (2.0 * (Q(pi + 1, pj) + Q(pi, pj - 1)) + (Q(pi, pj) + Q(pi + 1, pj - 1))) * C1_24;
        };
        auto D = [&](int pi, int pj) {
            return SCHEME == 2 ? ((Q(pi, pj) + Q(pi + 1, pj - 1)) + Q(pi, pj - 1)) * C1_12
                               : ((Q(pi + 1, pj) + Q(pi, pj - 1)) + 2.0 * (Q(pi, pj) + Q(pi + 1, pj - 1))) * C1_24;
        };
        auto B = [&](int pi, int pj) {
            return SCHEME == 2 ? (Q(pi, pj) + (Q(pi - 1, pj) + Q(pi, pj - 1))) * C1_12
                               : ((Q(pi, pj) + Q(pi - 1, pj - 1)) + 2.0 * (Q(pi - 1, pj) + Q(pi, pj - 1))) * C1_24;
        };
        auto Cc = [&](int pi, int pj) {
            return SCHEME == 2 ? ((Q(pi, pj) + Q(pi - 1, pj - 1)) + Q(pi, pj - 1)) * C1_12
                               : (2.0 * (Q(pi, pj) + Q(pi - 1, pj - 1)) + (Q(pi - 1, pj) + Q(pi, pj - 1))) * C1_24;
        };
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            if (i > ie || j > je) continue;
            const int li = tx + 1, lj = ty + 4 * r + 1;
            const long long c = IX3(d, i, j, k);
            const double KEc = s_KE[lj][li];
            if (i <= ie - 1) {
                const double KEx = (s_KE[lj][li + 1] - KEc) * g.IdxCu;
                if (SCHEME == 1) {   // Sadourny (1975) energy conserving (:206-216)
                    CAu[c] = 0.25 * ((Q(li, lj) * (s_vh[lj][li + 1] + s_vh[lj][li])) +
                                     (Q(li, lj - 1) * (s_vh[lj - 1][li] + s_vh[lj - 1][li + 1]))) *
                                 g.IdxCu - KEx;
                } else {             // (:229-237)
                    CAu[c] = (((A(li, lj) * s_vh[lj][li + 1]) + (Cc(li, lj) * s_vh[lj - 1][li])) +
                              ((B(li, lj) * s_vh[lj][li]) + (D(li, lj) * s_vh[lj - 1][li + 1]))) *
                                 g.IdxCu - KEx;
                }
            }
            if (j <= je - 1) {
                const double KEy = (s_KE[lj + 1][li] - KEc) * g.IdyCv;
                if (SCHEME == 1) {   // (:218-226)
                    CAv[c] = -(0.25 * ((Q(li - 1, lj) * (s_uh[lj][li - 1] + s_uh[lj + 1][li - 1])) +
                                       (Q(li, lj) * (s_uh[lj][li] + s_uh[lj + 1][li]))) *
                               g.IdyCv) - KEy;
                } else {             // (:239-247)
                    CAv[c] = -((((A(li - 1, lj) * s_uh[lj][li - 1]) + (Cc(li, lj + 1) * s_uh[lj + 1][li])) +
                                ((B(li, lj) * s_uh[lj][li]) + (D(li - 1, lj + 1) * s_uh[lj + 1][li - 1]))) *
                               g.IdyCv) - KEy;
                }
            }
        }
        __syncthreads();
    }
}

template <int SCHEME>
static void launch_cor_tile(const mom6h_coriolis_cs *cs, const CorUniform &g, const double *u, const double *v,
                            const double *h, const double *uh, const double *vh, double *CAu, double *CAv,
                            hipStream_t s)
{
    const Dims &d = cs->d;
    constexpr int TY = 8;
    const int nkz = (d.nk + 14) / 15;
    const dim3 grid((unsigned)((d.iec - d.isc + CT_X) / CT_X), (unsigned)((d.jec - d.jsc + TY) / TY), (unsigned)nkz);
    k_cor_tile<SCHEME, TY><<<grid, 256, 0, s>>>(d, g, u, v, h, uh, vh, cs->CoriolisBu, cs->mask2dBu, CAu, CAv);
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
    {   // constant metrics over the points the stencil reads
        const int ilo = d.isc - 2, ihi = d.iec + 2, jlo = d.jsc - 2, jhi = d.jec + 2;
        bool ok = true;
        for (auto f : {std::make_pair(areaT, &cs->u_areaT), std::make_pair(IareaBu, &cs->u_IareaBu),
                       std::make_pair(dyCv, &cs->u_dyCv), std::make_pair(dxCu, &cs->u_dxCu),
                       std::make_pair(dyCu, &cs->u_dyCu), std::make_pair(dxCv, &cs->u_dxCv),
                       std::make_pair(IdxCu, &cs->u_IdxCu), std::make_pair(IdyCv, &cs->u_IdyCv)})
            ok = ok && uniform_over(d, f.first, ilo, ihi, jlo, jhi, f.second);
        ok = ok && uniform_over(d, aq, d.isc - 1, d.iec, d.jsc - 1, d.jec, &cs->u_Area_q);
        cs->uniform = ok;
    }
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
    if (cs->uniform && !cs->general) {
        const CorUniform g{cs->u_areaT, cs->u_IareaBu, cs->u_dyCv, cs->u_dxCu, cs->u_dyCu,
                           cs->u_dxCv, cs->u_IdxCu, cs->u_IdyCv, cs->u_Area_q};
        if (cs->scheme == 1) launch_cor_tile<1>(cs, g, u, v, h, uh, vh, CAu, CAv, s);
        else if (cs->scheme == 2) launch_cor_tile<2>(cs, g, u, v, h, uh, vh, CAu, CAv, s);
        else launch_cor_tile<3>(cs, g, u, v, h, uh, vh, CAu, CAv, s);
        MOM6H_LAUNCH_CHECK();
        return MOM6H_OK;
    }
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
