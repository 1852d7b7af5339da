// mom6h_fused.hip -- fused momentum stencil for grids with uniform metrics.
//
// One launch performs, for every layer, what the reference does in four
// module calls (app/rk2_serial_driver.F90:236-270 and :316-350):
//   compute_transports (:544-561) -> hor_visc (src/serial/mom6_hor_visc_serial.F90,
//   Laplacian) -> CorAdCalc (src/serial/mom6_coriolis_serial.F90, Sadourny
//   energy scheme) -> velocity update  ou = base + fdt*(CA + diff)
// The transports, tendencies and potential vorticity never leave the chip: a
// block stages u, v, h of its tile plus a one-cell ring in LDS for one layer,
// forms q and KE on the tile's q/h points, then every thread finishes its
// velocity points.  It walks all layers, so the 2D masks and Coriolis
// parameter stay in L1/L2 for the whole layer walk.
//
// On a uniform grid every metric array holds one value; the kernels take those
// values (read from the arrays at set-up) and keep the reference's operation
// order, so the results are bit-identical to the general kernels.  The
// Kh/Ah ring of hor_visc and the never-written hArea rows of CorAdCalc only
// matter for diffu/diffv/CAu/CAv outside the velocity-update ranges; inside
// them the q values on row js-1 and column is-1 still see the zero hArea the
// reference reads, exactly as here.
#include "mom6h_fused.h"

namespace mom6h {

constexpr int MT_X = 64, MT_Y = 8;             // output tile
constexpr int MW = MT_X + 2, MH = MT_Y + 2;    // staged region [i0-1, i1+1] x [j0-1, j1+1]

__global__ void __launch_bounds__(256) k_momentum(MomArgs a)
{
    __shared__ double s_u[MH][MW], s_v[MH][MW], s_h[MH][MW];
    __shared__ double s_q[MH][MW], s_KE[MH][MW];
    const Dims &d = a.d;
    const UMetrics &g = a.g;
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int i0 = is + (tile % gridDim.x) * MT_X, j0 = js + (tile / gridDim.x) * MT_Y;
    const int bi = i0 - 1, bj = j0 - 1;
    const int tid = threadIdx.x;
    const int tx = tid % MT_X, ty = tid / MT_X;          // 64 x 4 threads, two rows each

    // u, v, h of the staged region [i0-1, i1+1] x [j0-1, j1+1] of one layer go
    // through registers: the next layer's loads are issued before the current
    // layer is computed (NS points per thread), and so are the base velocities
    // of the thread's two output rows
    constexpr int NS = (MW * MH + 255) / 256;
    double pu[NS], pv[NS], ph[NS], pbu[2], pbv[2];
    auto load_layer = [&](int k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            const int li = t % MW, lj = t / MW, i = bi + li, j = bj + lj;
            pu[r] = 0.0; pv[r] = 0.0; ph[r] = 0.0;
            if (t < MW * MH && i <= d.ied && j <= d.jed) {
                const long long c = IX3(d, i, j, k);
                pu[r] = a.u[c]; pv[r] = a.v[c]; ph[r] = a.h[c];
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            pbu[r] = 0.0; pbv[r] = 0.0;
            if (i <= ie && j <= je) {
                const long long c = IX3(d, i, j, k);
                pbu[r] = a.bu[c]; pbv[r] = a.bv[c];
            }
        }
    };
    load_layer(1);
    for (int k = 1; k <= d.nk; ++k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            if (t < MW * MH) {
                const int li = t % MW, lj = t / MW;
                s_u[lj][li] = pu[r]; s_v[lj][li] = pv[r]; s_h[lj][li] = ph[r];
            }
        }
        const double bu_r[2] = {pbu[0], pbu[1]}, bv_r[2] = {pbv[0], pbv[1]};
        __syncthreads();
        if (k < d.nk) load_layer(k + 1);
        // q on q points [i0-1, i1] x [j0-1, j1] (mom6_coriolis_serial.F90:134-168)
        // and KE on h points [i0, i1+1] x [j0, j1+1] (:196-203)
        for (int t = tid; t < MW * MH; t += blockDim.x) {
            const int li = t % MW, lj = t / MW, i = bi + li, j = bj + lj;
            if (li < MW - 1 && lj < MH - 1 && i <= ie && j <= je) {
                const long long c2 = IX2(d, i, j);
                const double dvdx = (s_v[lj][li + 1] * g.dyCv) - (s_v[lj][li] * g.dyCv);
                const double dudy = (s_u[lj + 1][li] * g.dxCu) - (s_u[lj][li] * g.dxCu);
                const double hAu0 = (j < js) ? 0.0 : 0.5 * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj][li + 1]));
                const double hAu1 = 0.5 * ((g.areaT * s_h[lj + 1][li]) + (g.areaT * s_h[lj + 1][li + 1]));
                const double hAv0 = (i < is) ? 0.0 : 0.5 * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj + 1][li]));
                const double hAv1 = 0.5 * ((g.areaT * s_h[lj][li + 1]) + (g.areaT * s_h[lj + 1][li + 1]));
                const double rel_vort = a.mask2dBu[c2] * (dvdx - dudy) * g.IareaBu;
                const double abs_vort = a.CoriolisBu[c2] + rel_vort;
                const double hArea_q = (hAu0 + hAu1) + (hAv0 + hAv1);
                s_q[lj][li] = abs_vort * (g.Area_q / (hArea_q + 1.0e-20));
            }
            if (li >= 1 && lj >= 1 && i <= ie && j <= je) {
                const double uc = s_u[lj][li], uw = s_u[lj][li - 1], vc = s_v[lj][li], vs = s_v[lj - 1][li];
                s_KE[lj][li] = 0.25 * ((g.dyCu * (uc * uc) + g.dyCu * (uw * uw)) +
                                       (g.dxCv * (vc * vc) + g.dxCv * (vs * vs))) / g.areaT;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int r = 0; r < 2; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            if (i > ie || j > je) continue;
            const int li = i - bi, lj = j - bj;
            const long long c = IX3(d, i, j, k);
            // Laplacian stress pieces (mom6_hor_visc_serial.F90:503-547, 807-1053)
            auto str_xx = [&](int pi, int pj) -> double {   // LDS coords of an h point
                const double dudx = g.DY_dxT * ((g.IdyCu * s_u[pj][pi]) - (g.IdyCu * s_u[pj][pi - 1]));
                const double dvdy = g.DX_dyT * ((g.IdxCv * s_v[pj][pi]) - (g.IdxCv * s_v[pj - 1][pi]));
                return -(a.Kh * (dudx - dvdy)) * (s_h[pj][pi] * 1.0);
            };
            auto str_xy = [&](int pi, int pj, int I, int J) -> double {   // q point
                const double dvdx = g.DY_dxBu * ((s_v[pj][pi + 1] * g.IdyCv) - (s_v[pj][pi] * g.IdyCv));
                const double dudy = g.DX_dyBu * ((s_u[pj + 1][pi] * g.IdxCu) - (s_u[pj][pi] * g.IdxCu));
                const double m = a.mask2dBu[IX2(d, I, J)];
                const double sh = (a.no_slip ? (2.0 - m) : m) * (dvdx + dudy);
                const double hq = 0.25 * ((s_h[pj][pi] + s_h[pj + 1][pi + 1]) + (s_h[pj][pi + 1] + s_h[pj + 1][pi]));
                const double str = -(a.Kh * sh);
                return a.no_slip ? str * (hq * 1.0) : str * (hq * m * 1.0);
            };
            if (i <= ie - 1) {
                // diffu (:1059-1067)
                const double h_u = 0.5 * (a.mask2dT[IX2(d, i, j)] * s_h[lj][li] + a.mask2dT[IX2(d, i + 1, j)] * s_h[lj][li + 1]);
                const double diffu = ((g.IdxCu * ((g.dx2q * str_xy(li, lj - 1, i, j - 1)) - (g.dx2q * str_xy(li, lj, i, j))) +
                                       g.IdyCu * ((g.dy2h * str_xx(li, lj)) - (g.dy2h * str_xx(li + 1, lj)))) *
                                      g.IareaCu) / (h_u + a.h_neglect);
                // CAu, Sadourny (:208-216); vh from compute_transports
                auto vh = [&](int pi, int pj) { return s_v[pj][pi] * 0.5 * (s_h[pj][pi] + s_h[pj + 1][pi]) * g.dxCv; };
                const double KEx = (s_KE[lj][li + 1] - s_KE[lj][li]) * g.IdxCu;
                const double CAu = 0.25 * ((s_q[lj][li] * (vh(li + 1, lj) + vh(li, lj))) +
                                           (s_q[lj - 1][li] * (vh(li, lj - 1) + vh(li + 1, lj - 1)))) * g.IdxCu - KEx;
                a.ou[c] = bu_r[r] + a.fdt * (CAu + diffu);
            }
            if (j <= je - 1) {
                // diffv (:1070-1078)
                const double h_v = 0.5 * (a.mask2dT[IX2(d, i, j)] * s_h[lj][li] + a.mask2dT[IX2(d, i, j + 1)] * s_h[lj + 1][li]);
                const double diffv = ((g.IdyCv * ((g.dy2q * str_xy(li - 1, lj, i - 1, j)) - (g.dy2q * str_xy(li, lj, i, j))) -
                                       g.IdxCv * ((g.dx2h * str_xx(li, lj)) - (g.dx2h * str_xx(li, lj + 1)))) *
                                      g.IareaCv) / (h_v + a.h_neglect);
                // CAv, Sadourny (:218-226); uh from compute_transports
                auto uh = [&](int pi, int pj) { return s_u[pj][pi] * 0.5 * (s_h[pj][pi] + s_h[pj][pi + 1]) * g.dyCu; };
                const double KEy = (s_KE[lj + 1][li] - s_KE[lj][li]) * g.IdyCv;
                const double CAv = -(0.25 * ((s_q[lj][li - 1] * (uh(li - 1, lj) + uh(li - 1, lj + 1))) +
                                             (s_q[lj][li] * (uh(li, lj) + uh(li, lj + 1)))) * g.IdyCv) - KEy;
                a.ov[c] = bv_r[r] + a.fdt * (CAv + diffv);
            }
        }
        __syncthreads();
    }
}

int launch_momentum(const MomArgs &a, hipStream_t s)
{
    const Dims &d = a.d;
    dim3 grid((d.iec - d.isc + 1 + MT_X - 1) / MT_X, (d.jec - d.jsc + 1 + MT_Y - 1) / MT_Y), block(256);
    k_momentum<<<grid, block, 0, s>>>(a);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

}  // namespace mom6h
