// mom6h_barotropic.hip -- the barotropic substep loop on gfx950.
//
// Arithmetic of src/serial/mom6_barotropic_serial.F90:178-356 (btstep,
// update_u, update_v) and, for periodic grids, the split API's extended
// transport ranges (btstep_do_step, :379-474) with a refill of the ubt, vbt and
// eta halos after every substep but the last (app/rk2_mpi_serial_driver.F90:
// 297-306).  The reference CUDA-C path issues ~9 dependent launches per
// substep from the host; here each substep is ONE launch of an LDS-tiled
// kernel that runs the whole substep (eta predictor, pressure forces, both
// Coriolis/velocity updates in the alternating order, transports, free
// surface, time averages, periodic refill) on its tile.  Transports that the
// reference never writes (uhbt on I = is-1 and ie, vhbt on J = js-1 and je
// with the narrow ranges) read as zero, as they do there.
#include <algorithm>
#include <mutex>
#include <utility>
#include <type_traits>
#include <vector>

#include "mom6h_internal.h"

namespace mom6h {

// ---------------------------------------------------------------------------
// One substep per launch, 2D tiles staged in LDS.  A block owns the output
// tile T = [i0, i0+TX-1] x [j0, j0+TY-1] and recomputes the few cells around
// it that its stencils need (radius 2 to the west/south, 1 to the east/north),
// so no grid-wide barrier is needed: consecutive launches on one stream order
// the substeps.  State ping-pongs between two buffer sets; the time averages
// are accumulated in place.
constexpr int BT_TX = 64;
constexpr int BT_W = BT_TX + 3;

__global__ void k_bt_copy_in(Dims d, const double *eta_in, const double *ubt_in, const double *vbt_in,
                             double *eta, double *ubt, double *vbt, double *ubt_av, double *vbt_av,
                             double *eta_b, double *ubt_b, double *vbt_b)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.s2) return;
    eta[q] = eta_in[q];
    ubt[q] = ubt_in[q];
    vbt[q] = vbt_in[q];
    ubt_av[q] = 0.0;
    vbt_av[q] = 0.0;
    // the second state set starts at zero (a control structure's is zeroed
    // once at init; scratch shared between calls is cleared on every call)
    if (eta_b) {
        eta_b[q] = 0.0;
        ubt_b[q] = 0.0;
        vbt_b[q] = 0.0;
    }
}

// Output tile BT_TX x BT_TY (BT_TY = 4, 8 or 16 rows: the launch's choice),
// 256 threads, BT_TY/4 outputs per thread.  Three LDS arrays hold eta, ubt
// and vbt of the staged region and are updated in place: eta becomes the
// predicted eta (each point reads only its own eta), each velocity update
// reads the other component and its own old value only, and the old values
// the transports of a thread's outputs need are kept in registers before the
// first update.  The pressure forces are recomputed from the predicted eta
// where each update needs them (same operands, same values).  At 8 rows the
// block needs 17.7 KB of LDS, and the whole 720x720 grid fits the chip in
// one round of blocks.
template <typename M, int BT_TY>
__device__ __forceinline__ void bt_body(const BT &p, const M &m, const BTStep &st, int tile_x, int tile_y)
{
    constexpr int BT_H = BT_TY + 3, NOUT = BT_TY / 4;
    constexpr bool UNI = std::is_same<M, BTUniform>::value;   // Datv and f4 staged with the state
    __shared__ double s_e[BT_H][BT_W], s_u[BT_H][BT_W], s_v[BT_H][BT_W];
    __shared__ double s_dv[UNI ? BT_H : 1][UNI ? BT_W : 1], s_f4[UNI ? BT_H : 1][UNI ? BT_W : 1];
    const Dims &d = p.d;
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const int i0 = d.isd + tile_x * BT_TX, j0 = d.jsd + tile_y * BT_TY;
    const int i1 = i0 + BT_TX - 1, j1 = j0 + BT_TY - 1;
    const int bi = i0 - 2, bj = j0 - 2;  // LDS origin
    const int tid = threadIdx.x;
    const bool v_first = ((st.n + p.first_direction) % 2) == 1;
    const double dtbt = p.dtbt;

    // L: old state on [i0-2, i1+1] x [j0-2, j1+1]
    for (int t = tid; t < BT_W * BT_H; t += 256) {
        const int li = t % BT_W, lj = t / BT_W, i = bi + li, j = bj + lj;
        double e = 0.0, uu = 0.0, vv = 0.0, dv = 0.0, f4 = 0.0;
        if (i >= d.isd && i <= d.ied && j >= d.jsd && j <= d.jed) {
            const long long c = IX2(d, i, j);
            e = st.eta_o[c]; uu = st.ubt_o[c]; vv = st.vbt_o[c];
            if (UNI) { dv = m.datv(c); f4 = m.f4u(1, c); }
        }
        s_e[lj][li] = e; s_u[lj][li] = uu; s_v[lj][li] = vv;
        if (UNI) { s_dv[UNI ? lj : 0][UNI ? li : 0] = dv; s_f4[UNI ? lj : 0][UNI ? li : 0] = f4; }
    }
    // the running averages of this thread's outputs
    double ua_r[NOUT], va_r[NOUT];
#pragma unroll
    for (int r = 0; r < NOUT; ++r) {
        const int t = tid + 256 * r;
        const int i = i0 + t % BT_TX, j = j0 + t / BT_TX;
        ua_r[r] = 0.0; va_r[r] = 0.0;
        if (i <= d.ied && j <= d.jed) {
            ua_r[r] = p.ubt_av[IX2(d, i, j)];
            va_r[r] = p.vbt_av[IX2(d, i, j)];
        }
    }
    __syncthreads();
    // coefficient access: staged values (uniform grids) or the arrays
    auto DV = [&](int lj, int li, long long c) -> double {
        if constexpr (UNI) return s_dv[lj][li];
        else return m.datv(c);
    };
    auto F4U = [&](int comp, int lj, int li, long long c) -> double {
        if constexpr (UNI) return s_f4[lj][li];
        else return m.f4u(comp, c);
    };
    auto F4V = [&](int comp, int lj, int li, long long c) -> double {
        if constexpr (UNI) return s_f4[lj][li];
        else return m.f4v(comp, c);
    };
    // old values the transports of this thread's outputs need
    double e_old[NOUT], u_old[NOUT], uw_old[NOUT], v_old[NOUT], vs_old[NOUT];
#pragma unroll
    for (int r = 0; r < NOUT; ++r) {
        const int t = tid + 256 * r;
        const int li = 2 + t % BT_TX, lj = 2 + t / BT_TX;
        e_old[r] = s_e[lj][li];
        u_old[r] = s_u[lj][li]; uw_old[r] = s_u[lj][li - 1];
        v_old[r] = s_v[lj][li]; vs_old[r] = s_v[lj - 1][li];
    }
    __syncthreads();
    // P1: eta predictor (:224-231) on [i0-1, i1+1] x [j0-1, j1+1], in place
    for (int t = tid; t < BT_W * BT_H; t += 256) {
        const int li = t % BT_W, lj = t / BT_W, i = bi + li, j = bj + lj;
        if (i < i0 - 1 || j < j0 - 1 || i < is || i > ie || j < js || j > je) continue;
        const long long c = IX2(d, i, j);
        s_e[lj][li] = s_e[lj][li] + (dtbt * m.iareaT(c)) *
            (((m.datu(IX2(d, i - 1, j)) * s_u[lj][li - 1]) - (m.datu(c) * s_u[lj][li])) +
             ((DV(lj - 1, li, IX2(d, i, j - 1)) * s_v[lj - 1][li]) - (DV(lj, li, c) * s_v[lj][li])));
    }
    __syncthreads();
    // pressure forces (:234-247) from the predicted eta
    auto pf_u = [&](int lj, int li, int i, int j, long long c) -> double {
        return ((s_e[lj][li] * m.gE(c)) - (s_e[lj][li + 1] * m.gW(IX2(d, i + 1, j)))) * p.dgeo_de * m.idxCu(c);
    };
    auto pf_v = [&](int lj, int li, int i, int j, long long c) -> double {
        return ((s_e[lj][li] * m.gN(c)) - (s_e[lj + 1][li] * m.gS(IX2(d, i, j + 1)))) * p.dgeo_de * m.idyCv(c);
    };
    // the velocity updates (:250-260), with the other component as it stands in LDS
    auto new_v = [&](int lj, int li, int i, int j, long long c) -> double {
        double vn = s_v[lj][li];
        if (i >= is && i <= ie && j >= js && j <= je - 1) {
            const double Cor = -1.0 * (((F4V(1, lj, li, c) * s_u[lj][li - 1]) + (F4V(4, lj, li, c) * s_u[lj + 1][li])) +
                                       ((F4V(2, lj, li, c) * s_u[lj][li]) + (F4V(3, lj, li, c) * s_u[lj + 1][li - 1])));
            vn = m.remv(c) * (vn + dtbt * (Cor + pf_v(lj, li, i, j, c)));
        }
        return vn;
    };
    auto new_u = [&](int lj, int li, int i, int j, long long c) -> double {
        double un = s_u[lj][li];
        if (i >= is && i <= ie - 1 && j >= js && j <= je) {
            const double Cor = (((F4U(4, lj, li, c) * s_v[lj][li + 1]) + (F4U(1, lj, li, c) * s_v[lj - 1][li])) +
                                ((F4U(3, lj, li, c) * s_v[lj][li]) + (F4U(2, lj, li, c) * s_v[lj - 1][li + 1])));
            un = m.remu(c) * (un + dtbt * (Cor + pf_u(lj, li, i, j, c)));
        }
        return un;
    };
    // P2: the first update -- vbt on [i0-1, i1+1] x [j0-1, j1] or ubt on [i0-1, i1] x [j0-1, j1+1]
    for (int t = tid; t < BT_W * BT_H; t += 256) {
        const int li = t % BT_W, lj = t / BT_W, i = bi + li, j = bj + lj;
        if (i < i0 - 1 || j < j0 - 1) continue;
        const long long c = IX2(d, i, j);
        if (v_first) {
            if (j <= j1) s_v[lj][li] = new_v(lj, li, i, j, c);
        } else {
            if (i <= i1) s_u[lj][li] = new_u(lj, li, i, j, c);
        }
    }
    __syncthreads();
    // P3: the second update -- ubt on [i0-1, i1] x [j0, j1] or vbt on [i0, i1] x [j0-1, j1]
    for (int t = tid; t < BT_W * BT_H; t += 256) {
        const int li = t % BT_W, lj = t / BT_W, i = bi + li, j = bj + lj;
        if (i < i0 - 1 || j < j0 - 1 || i > i1 || j > j1) continue;
        const long long c = IX2(d, i, j);
        if (v_first) {
            if (j >= j0) s_u[lj][li] = new_u(lj, li, i, j, c);
        } else {
            if (i >= i0) s_v[lj][li] = new_v(lj, li, i, j, c);
        }
    }
    __syncthreads();
    // P4: transports (:262-272), free surface (:274-279), averages (:281-291), output
    const double trans_wt1 = 1.0 + p.bebt, trans_wt2 = -p.bebt;
    const double inv_nstep = 1.0 / (double)p.nstep;
    const int iu0 = p.extended ? is - 1 : is, iu1 = p.extended ? ie : ie - 1;
    const int jv0 = p.extended ? js - 1 : js, jv1 = p.extended ? je : je - 1;
    const bool refill = (p.px || p.py) && st.n < p.nstep;
    const bool last = st.n == p.nstep;
    const int P = ie - is + 1, Q = je - js + 1;
    auto uhbt = [&](int i, int j, double un, double uo) -> double {
        if (i < iu0 || i > iu1 || j < js || j > je) return 0.0;
        return m.datu(IX2(d, i, j)) * (trans_wt1 * un + trans_wt2 * uo);
    };
    auto vhbt = [&](int i, int j, int li, int lj, double vn, double vo) -> double {
        if (i < is || i > ie || j < jv0 || j > jv1) return 0.0;
        return DV(lj, li, IX2(d, i, j)) * (trans_wt1 * vn + trans_wt2 * vo);
    };
#pragma unroll
    for (int r = 0; r < NOUT; ++r) {
        const int t = tid + 256 * r;
        const int i = i0 + t % BT_TX, j = j0 + t / BT_TX;
        if (i > d.ied || j > d.jed) continue;
        const int li = i - bi, lj = j - bj;
        const long long c = IX2(d, i, j);
        const bool in_eta = i >= is && i <= ie && j >= js && j <= je;
        const bool in_u = i >= is && i <= ie - 1 && j >= js && j <= je;
        const bool in_v = i >= is && i <= ie && j >= js && j <= je - 1;
        // ubt/vbt carry their new value only on the update ranges (the
        // updates leave the old value elsewhere)
        const double un = s_u[lj][li], vn = s_v[lj][li];
        double e = e_old[r];
        if (in_eta)
            e = e - dtbt * m.iareaT(c) * ((uhbt(i, j, un, u_old[r]) - uhbt(i - 1, j, s_u[lj][li - 1], uw_old[r])) +
                                          (vhbt(i, j, li, lj, vn, v_old[r]) -
                                           vhbt(i, j - 1, li, lj - 1, s_v[lj - 1][li], vs_old[r])));
        double ua = ua_r[r], va = va_r[r];
        if (in_u) { ua = ua + un * inv_nstep; p.ubt_av[c] = ua; }
        if (in_v) { va = va + vn * inv_nstep; p.vbt_av[c] = va; }
        const bool halo_cell = i < is || i > ie || j < js || j > je;
        if (!(refill && halo_cell && (p.px ? true : (i >= is && i <= ie)) && (p.py ? true : (j >= js && j <= je)) &&
              ((p.px && (i < is || i > ie)) || (p.py && (j < js || j > je))))) {
            st.eta_n[c] = e; st.ubt_n[c] = un; st.vbt_n[c] = vn;
        }
        if (refill && !halo_cell) {
            // write the periodic images of this compute cell (halo refill, :301-303)
            const int ni_[3] = {i, (p.px && i > ie - (is - d.isd)) ? i - P : ((p.px && i < is + (d.ied - ie)) ? i + P : -1), 0};
            const int nj_[3] = {j, (p.py && j > je - (js - d.jsd)) ? j - Q : ((p.py && j < js + (d.jed - je)) ? j + Q : -1), 0};
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    if (a == 0 && b == 0) continue;
                    const int ti = ni_[a], tj = nj_[b];
                    if (ti < d.isd || ti > d.ied || tj < d.jsd || tj > d.jed) continue;
                    const long long ct = IX2(d, ti, tj);
                    st.eta_n[ct] = e; st.ubt_n[ct] = un; st.vbt_n[ct] = vn;
                }
        }
        if (last) {  // outputs (:296-302)
            st.out_u[c] = ua;
            st.out_v[c] = va;
            st.out_eta[c] = e;
        }
    }
}

template <typename M, int TY>
__global__ void __launch_bounds__(256) k_bt_substep(BT p, M m, BTStep st)
{
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bt_body<M, TY>(p, m, st, tile % gridDim.x, tile / gridDim.x);
}

// ---------------------------------------------------------------------------
// Several substeps per launch (temporal blocking) for the substeps between the
// first and the last of a call, on uniform grids.  The reference refills the
// periodic halos after every substep but the last, so from the second substep
// on a halo cell always holds the value of its periodic image: the substeps in
// between run on a torus.  A block therefore stages the region its output tile
// needs for ns substeps -- radius 2 to the west/south and 1 to the east/north
// per substep -- and maps every staged position to its canonical cell (the
// periodic image inside the compute domain; a position outside the arrays in
// a wall direction holds zero and is never updated).  Each position is updated
// as its canonical cell, with the state, Datv and f4 of the neighbouring
// staged positions (the refilled images; Datv's halo rows equal their images,
// which the host checks before it uses this kernel).  The region that holds
// valid values shrinks by the stencil radius each substep and is the output
// tile after the last.  The output tile is written whole, halo positions
// included (their images), so the state after the launch is what the reference
// holds after its refill.  The first substep of a call (which reads the halos
// as stored) and the last (after which nothing is refilled) stay with the
// one-substep kernel.
struct BTMulti {
    const double *eta_o, *ubt_o, *vbt_o;   // state before substep n0
    double *eta_n, *ubt_n, *vbt_n;         // state after substep n0 + ns - 1
    int n0, ns;                            // first substep, number of substeps (1 .. TS)
};

// M(w) = ceil(2^20 / w) for widths w < 256, so that t / w = (t * M(w)) >> 20
// exactly for t < 4096; read from constant memory with a uniform index (a
// scalar load) instead of an integer division per call.
struct MulShift {
    unsigned v[256];
};
constexpr MulShift make_mulshift()
{
    MulShift t{};
    for (unsigned w = 1; w < 256; ++w) t.v[w] = ((1u << 20) + w - 1) / w;
    return t;
}
__constant__ MulShift kMulShift = make_mulshift();

// Calls f(x0 + t % w, y0 + t / w) for the positions t < w*h of a rectangle of
// width w >= 8 and at most NT*MAXI positions (<= 4096: t / w as an exact
// multiply-shift), t spread over the block's NT threads.  The products fit in
// 24-bit operands (t < 2^12, M <= 2^17, q < 2^9, w < 2^12), so they take the
// full-rate 24-bit multiply rather than the quarter-rate 32-bit one.
template <int NT, int MAXI, typename F>
__device__ __forceinline__ void rect_each(int x0, int x1, int y0, int y1, int tid, F &&f)
{
    static_assert(NT * MAXI <= 4096, "positions beyond the exact multiply-shift range");
    const unsigned w = (unsigned)(x1 - x0 + 1), n = w * (unsigned)(y1 - y0 + 1);
    const unsigned M = kMulShift.v[w];   // 8 <= w < 256
#pragma unroll
    for (int r = 0; r < MAXI; ++r) {
        const unsigned t = (unsigned)tid + (unsigned)NT * r;
        if (t < n) {
            const unsigned q = __umul24(t, M) >> 20;
            const int li = x0 + (int)(t - __umul24(q, w)), lj = y0 + (int)q;
            // staged positions: small and non-negative, so that their LDS
            // offsets take 24-bit multiplies
            __builtin_assume(li >= 0 && li < 4096 && lj >= 0 && lj < 4096);
            f(li, lj);
        }
    }
}

// ROW: Datv and f4 constant along each row of the compute domain (walls close
// whole rows, f depends on latitude only; the host checks it): one staged value
// per row, and the predicted eta goes to an array of its own (no barrier
// between reading the old values and the predictor).
template <int NT, int TX, int TY, int TS, bool ROW>
__device__ __forceinline__ void bt_multi_body(const BT &p, const BTUniform &m, const BTMulti &st, int tile_x,
                                              int tile_y)
{
    constexpr int RW = TX + 3 * TS, RH = TY + 3 * TS;   // staged region (LDS)
    constexpr int NOUT = (TX * TY + NT - 1) / NT;       // output points per thread (t < TX*TY)
    constexpr int NR = ((TX + 3 * (TS - 1)) * (TY + 3 * (TS - 1)) + NT - 1) / NT;          // valid region
    constexpr int NP = ((TX + 3 * (TS - 1) + 2) * (TY + 3 * (TS - 1) + 2) + NT - 1) / NT;  // update regions
    static_assert(RW * RH <= 4096 && RW + 2 < 256, "staged region too large for the multiply-shift index");
    __shared__ double s_e[RH][RW], s_u[RH][RW], s_v[RH][RW];
    __shared__ double s_dv[RH][ROW ? 1 : RW], s_f4[RH][ROW ? 1 : RW];
    __shared__ double s_ep[ROW ? RH : 1][ROW ? RW : 1];   // ROW: the predicted eta
    __shared__ int s_cx[RW], s_cy[RH];
    // range tests of each staged column (x) and row (y) on its canonical
    // index, one bit each, formed once per launch: bit 0 inside the compute
    // domain (is..ie / js..je), bit 1 inside it but for the last face
    // (..ie-1 / ..je-1: the u / v update ranges), bit 2 inside the u / v
    // transport range (iu0..iu1 / jv0..jv1), bit 3 the same for the index - 1
    __shared__ int s_fx[RW], s_fy[RH];
    auto DV = [&](int lj, int li) -> double { return s_dv[lj][ROW ? 0 : li]; };
    auto F4 = [&](int lj, int li) -> double { return s_f4[lj][ROW ? 0 : li]; };
    auto EP = [&](int lj, int li) -> double { return ROW ? s_ep[ROW ? lj : 0][ROW ? li : 0] : s_e[lj][li]; };
    const Dims &d = p.d;
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const int i0 = d.isd + tile_x * TX, j0 = d.jsd + tile_y * TY;
    const int bi = i0 - 2 * TS, bj = j0 - 2 * TS;   // LDS origin
    const int tid = threadIdx.x;
    const double dtbt = p.dtbt;
    const int iu0 = p.extended ? is - 1 : is, iu1 = p.extended ? ie : ie - 1;
    const int jv0 = p.extended ? js - 1 : js, jv1 = p.extended ? je : je - 1;
    auto flags = [](int c, int lo, int hi, int t0, int t1) -> int {
        return ((c >= lo && c <= hi) ? 1 : 0) | ((c >= lo && c <= hi - 1) ? 2 : 0) |
               ((c >= t0 && c <= t1) ? 4 : 0) | ((c - 1 >= t0 && c - 1 <= t1) ? 8 : 0);
    };
    // canonical index of each staged column and row, and its range bits
    for (int t = tid; t < RW + RH; t += NT) {
        if (t < RW) {
            const int x = bi + t, P = ie - is + 1;
            const int c = p.px ? is + ((x - is) % P + P) % P : x;
            s_cx[t] = c;
            s_fx[t] = flags(c, is, ie, iu0, iu1);
        } else {
            const int y = bj + (t - RW), Q = je - js + 1;
            const int c = p.py ? js + ((y - js) % Q + Q) % Q : y;
            s_cy[t - RW] = c;
            s_fy[t - RW] = flags(c, js, je, jv0, jv1);
        }
    }
    __syncthreads();
    rect_each<NT, (RW * RH + NT - 1) / NT>(0, RW - 1, 0, RH - 1, tid, [&](int li, int lj) {
        const int i = s_cx[li], j = s_cy[lj];
        double e = 0.0, uu = 0.0, vv = 0.0, dv = 0.0, f4 = 0.0;
        if (i >= d.isd && i <= d.ied && j >= d.jsd && j <= d.jed) {
            const long long c = IX2(d, i, j);
            e = st.eta_o[c]; uu = st.ubt_o[c]; vv = st.vbt_o[c];
            if (!ROW) { dv = m.Datv[c]; f4 = m.f4[c]; }
        }
        s_e[lj][li] = e; s_u[lj][li] = uu; s_v[lj][li] = vv;
        if (!ROW) { s_dv[lj][ROW ? 0 : li] = dv; s_f4[lj][ROW ? 0 : li] = f4; }
    });
    if (ROW) {
        for (int t = tid; t < RH; t += NT) {
            const int j = s_cy[t];
            const bool in = j >= d.jsd && j <= d.jed;
            s_dv[t][0] = in ? m.Datv_row[j - d.jsd] : 0.0;
            s_f4[t][0] = in ? m.f4_row[j - d.jsd] : 0.0;
        }
    }
    // running averages of this thread's output points (the array positions)
    double ua_r[NOUT], va_r[NOUT];
#pragma unroll
    for (int r = 0; r < NOUT; ++r) {
        const int t = tid + NT * r;
        const int i = i0 + t % TX, j = j0 + t / TX;
        ua_r[r] = 0.0; va_r[r] = 0.0;
        if (t < TX * TY && i <= d.ied && j <= d.jed) {
            ua_r[r] = p.ubt_av[IX2(d, i, j)];
            va_r[r] = p.vbt_av[IX2(d, i, j)];
        }
    }
    __syncthreads();
    // pressure forces and the velocity updates (:234-260) of the canonical
    // cell (i, j) at staged position (li, lj); the coefficients are uniform
    // but for Datv and f4
    auto new_v = [&](int lj, int li) -> double {
        double vn = s_v[lj][li];
        if ((s_fx[li] & 1) && (s_fy[lj] & 2)) {   // is..ie x js..je-1
            const double f4 = F4(lj, li);
            const double Cor = -1.0 * (((f4 * s_u[lj][li - 1]) + (f4 * s_u[lj + 1][li])) +
                                       ((f4 * s_u[lj][li]) + (f4 * s_u[lj + 1][li - 1])));
            const double pf = ((EP(lj, li) * m.gtot) - (EP(lj + 1, li) * m.gtot)) * p.dgeo_de * m.IdyCv;
            vn = m.bt_rem_v * (vn + dtbt * (Cor + pf));
        }
        return vn;
    };
    auto new_u = [&](int lj, int li) -> double {
        double un = s_u[lj][li];
        if ((s_fx[li] & 2) && (s_fy[lj] & 1)) {   // is..ie-1 x js..je
            const double f4 = F4(lj, li);
            const double Cor = (((f4 * s_v[lj][li + 1]) + (f4 * s_v[lj - 1][li])) +
                                ((f4 * s_v[lj][li]) + (f4 * s_v[lj - 1][li + 1])));
            const double pf = ((EP(lj, li) * m.gtot) - (EP(lj, li + 1) * m.gtot)) * p.dgeo_de * m.IdxCu;
            un = m.bt_rem_u * (un + dtbt * (Cor + pf));
        }
        return un;
    };
    const double trans_wt1 = 1.0 + p.bebt, trans_wt2 = -p.bebt;
    const double inv_nstep = 1.0 / (double)p.nstep;
    for (int s = 0; s < st.ns; ++s) {
        const int n = st.n0 + s;
        const int k = st.ns - 1 - s;   // substeps of this launch after this one
        // R = [A0, A1] x [B0, B1]: the positions this substep leaves valid
        const int A0 = 2 * TS - 2 * k, A1 = 2 * TS + TX - 1 + k;
        const int B0 = 2 * TS - 2 * k, B1 = 2 * TS + TY - 1 + k;
        const bool v_first = ((n + p.first_direction) % 2) == 1;
        // this thread's points of R: t = tid + NT r, (A0 + t % rw, B0 + t / rw)
        const unsigned rw = (unsigned)(A1 - A0 + 1), nR = rw * (unsigned)(B1 - B0 + 1);
        const unsigned Mr = kMulShift.v[rw];
        auto r_pos = [&](int r, int &li, int &lj) -> bool {
            const unsigned t = (unsigned)tid + (unsigned)NT * r, q = __umul24(t, Mr) >> 20;
            li = A0 + (int)(t - __umul24(q, rw));
            lj = B0 + (int)q;
            __builtin_assume(li >= 0 && li < 4096 && lj >= 0 && lj < 4096);
            return t < nR;
        };
        // old values the transports of these points need
        double e_old[NR], u_old[NR], uw_old[NR], v_old[NR], vs_old[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            int li, lj;
            if (r_pos(r, li, lj)) {
                e_old[r] = s_e[lj][li];
                u_old[r] = s_u[lj][li]; uw_old[r] = s_u[lj][li - 1];
                v_old[r] = s_v[lj][li]; vs_old[r] = s_v[lj - 1][li];
            }
        }
        if (!ROW) __syncthreads();   // the predictor below overwrites s_e in place
        // eta predictor (:224-231): in place, or (ROW) into s_ep, which then
        // holds eta itself where no prediction applies
        rect_each<NT, NP>(A0 - 1, A1 + 1, B0 - 1, B1 + 1, tid, [&](int li, int lj) {
            const bool in = (s_fx[li] & s_fy[lj] & 1) != 0;
            if (!ROW && !in) return;
            const double e0 = s_e[lj][li];
            const double ep = in ? e0 + (dtbt * m.IareaT) *
                                            (((m.Datu * s_u[lj][li - 1]) - (m.Datu * s_u[lj][li])) +
                                             ((DV(lj - 1, li) * s_v[lj - 1][li]) - (DV(lj, li) * s_v[lj][li])))
                                 : e0;
            if (ROW) s_ep[ROW ? lj : 0][ROW ? li : 0] = ep;
            else s_e[lj][li] = ep;
        });
        __syncthreads();
        // the first velocity update
        if (v_first)
            rect_each<NT, NP>(A0 - 1, A1 + 1, B0 - 1, B1, tid, [&](int li, int lj) {
                s_v[lj][li] = new_v(lj, li);
            });
        else
            rect_each<NT, NP>(A0 - 1, A1, B0 - 1, B1 + 1, tid, [&](int li, int lj) {
                s_u[lj][li] = new_u(lj, li);
            });
        __syncthreads();
        // the second velocity update
        if (v_first)
            rect_each<NT, NP>(A0 - 1, A1, B0, B1, tid, [&](int li, int lj) {
                s_u[lj][li] = new_u(lj, li);
            });
        else
            rect_each<NT, NP>(A0, A1, B0 - 1, B1, tid, [&](int li, int lj) {
                s_v[lj][li] = new_v(lj, li);
            });
        __syncthreads();
        // time averages (:281-291) of the output points: the array faces
        // inside the update ranges
#pragma unroll
        for (int r = 0; r < NOUT; ++r) {
            const int t = tid + NT * r;
            const int i = i0 + t % TX, j = j0 + t / TX;
            const int li = 2 * TS + t % TX, lj = 2 * TS + t / TX;
            if (t >= TX * TY) continue;
            if (i >= is && i <= ie - 1 && j >= js && j <= je) ua_r[r] = ua_r[r] + s_u[lj][li] * inv_nstep;
            if (i >= is && i <= ie && j >= js && j <= je - 1) va_r[r] = va_r[r] + s_v[lj][li] * inv_nstep;
        }
        // transports (:262-272) and the free surface (:274-279) on R, in place
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            int li, lj;
            if (!r_pos(r, li, lj)) continue;
            const int fx = s_fx[li], fy = s_fy[lj];
            if (!(fx & fy & 1)) continue;   // outside is..ie x js..je
            const double un = s_u[lj][li], uw = s_u[lj][li - 1], vn = s_v[lj][li], vs = s_v[lj - 1][li];
            const double uh_e = (fx & 4) ? m.Datu * (trans_wt1 * un + trans_wt2 * u_old[r]) : 0.0;
            const double uh_w = (fx & 8) ? m.Datu * (trans_wt1 * uw + trans_wt2 * uw_old[r]) : 0.0;
            const double vh_n = (fy & 4) ? DV(lj, li) * (trans_wt1 * vn + trans_wt2 * v_old[r]) : 0.0;
            const double vh_s = (fy & 8) ? DV(lj - 1, li) * (trans_wt1 * vs + trans_wt2 * vs_old[r]) : 0.0;
            s_e[lj][li] = e_old[r] - dtbt * m.IareaT * ((uh_e - uh_w) + (vh_n - vh_s));
        }
        __syncthreads();
    }
    // the state after the launch on the output tile, halo positions included
#pragma unroll
    for (int r = 0; r < NOUT; ++r) {
        const int t = tid + NT * r;
        const int i = i0 + t % TX, j = j0 + t / TX;
        if (t >= TX * TY || i > d.ied || j > d.jed) continue;
        const int li = 2 * TS + t % TX, lj = 2 * TS + t / TX;
        const long long c = IX2(d, i, j);
        st.eta_n[c] = s_e[lj][li]; st.ubt_n[c] = s_u[lj][li]; st.vbt_n[c] = s_v[lj][li];
        if (i >= is && i <= ie - 1 && j >= js && j <= je) p.ubt_av[c] = ua_r[r];
        if (i >= is && i <= ie && j >= js && j <= je - 1) p.vbt_av[c] = va_r[r];
    }
}

// Tile and block of the blocked kernel for TS substeps per launch.  66 x 16
// (512 threads) covers a 720-wide grid and its halos in 11 x 46 tiles, one
// round of two blocks per CU; 4 substeps take 66 x 32 tiles of 1024 threads
// (one block per CU, 11 x 23 tiles), which recompute less of their halo.
template <int TS>
struct BTMShape {
    static constexpr int TX = 66, TY = 16, NT = 512;
};
template <>
struct BTMShape<4> {
    static constexpr int TX = 66, TY = 32, NT = 1024;
};
template <>
struct BTMShape<5> : BTMShape<4> {};

inline dim3 bt_multi_grid(const Dims &d, int ts)
{
    const int tx = ts >= 4 ? BTMShape<4>::TX : BTMShape<3>::TX, ty = ts >= 4 ? BTMShape<4>::TY : BTMShape<3>::TY;
    return dim3((d.n1 + tx - 1) / tx, (d.n2 + ty - 1) / ty);
}
inline int bt_multi_threads(int ts) { return ts >= 4 ? BTMShape<4>::NT : BTMShape<3>::NT; }

template <int TS, bool ROW>
__global__ void __launch_bounds__(BTMShape<TS>::NT) k_bt_multi(BT p, BTUniform m, BTMulti st)
{
    using S = BTMShape<TS>;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    bt_multi_body<S::NT, S::TX, S::TY, TS, ROW>(p, m, st, tile % gridDim.x, tile / gridDim.x);
}

template <bool ROW>
void launch_bt_multi(int ts, dim3 g, const BT &p, const BTUniform &m, const BTMulti &st, hipStream_t s)
{
    const dim3 b(bt_multi_threads(ts));
    if (ts == 2) k_bt_multi<2, ROW><<<g, b, 0, s>>>(p, m, st);
    else if (ts == 3) k_bt_multi<3, ROW><<<g, b, 0, s>>>(p, m, st);
    else if (ts == 4) k_bt_multi<4, ROW><<<g, b, 0, s>>>(p, m, st);
    else k_bt_multi<5, ROW><<<g, b, 0, s>>>(p, m, st);
}

// ---- the ensemble form: every member's substep n (or substeps n0 ..
// n0+ns-1, blocked) in one launch.  Member state ping-pongs between A (where
// the copy-in puts it) and B; from_a says which set a launch reads.
__global__ void k_bt_copy_in_batch(const BTMember *__restrict__ args, BatchMap mp)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const BTMember &a = args[m];
    const long long q = (long long)local * 256 + threadIdx.x;
    if (q >= a.p.d.s2) return;
    a.A[0][q] = a.in[0][q];
    a.A[1][q] = a.in[1][q];
    a.A[2][q] = a.in[2][q];
    a.p.ubt_av[q] = 0.0;
    a.p.vbt_av[q] = 0.0;
}

template <int TY>
__global__ void __launch_bounds__(256) k_bt_substep_batch(const BTMember *__restrict__ args, BatchMap mp, int n,
                                                          int from_a)
{
    int m, local;
    batch_locate(mp, xcd_tile(blockIdx.x, gridDim.x), m, local);
    const BTMember a = args[m];
    BTStep st;
    st.eta_o = from_a ? a.A[0] : a.B[0]; st.ubt_o = from_a ? a.A[1] : a.B[1]; st.vbt_o = from_a ? a.A[2] : a.B[2];
    st.eta_n = from_a ? a.B[0] : a.A[0]; st.ubt_n = from_a ? a.B[1] : a.A[1]; st.vbt_n = from_a ? a.B[2] : a.A[2];
    st.out_u = a.out[0]; st.out_v = a.out[1]; st.out_eta = a.out[2];
    st.n = n;
    bt_body<BTUniform, TY>(a.p, a.m, st, local % mp.nbx[m], local / mp.nbx[m]);
}

template <int TS, bool ROW>
__global__ void __launch_bounds__(BTMShape<TS>::NT) k_bt_multi_batch(const BTMember *__restrict__ args, BatchMap mp, int n0,
                                                            int ns, int from_a)
{
    int m, local;
    batch_locate(mp, xcd_tile(blockIdx.x, gridDim.x), m, local);
    const BTMember a = args[m];
    BTMulti st;
    st.eta_o = from_a ? a.A[0] : a.B[0]; st.ubt_o = from_a ? a.A[1] : a.B[1]; st.vbt_o = from_a ? a.A[2] : a.B[2];
    st.eta_n = from_a ? a.B[0] : a.A[0]; st.ubt_n = from_a ? a.B[1] : a.A[1]; st.vbt_n = from_a ? a.B[2] : a.A[2];
    st.n0 = n0;
    st.ns = ns;
    using S = BTMShape<TS>;
    bt_multi_body<S::NT, S::TX, S::TY, TS, ROW>(a.p, a.m, st, local % mp.nbx[m], local / mp.nbx[m]);
}

// btstep of every member (all with nstep substeps and uniform coefficients);
// substeps 2 .. nstep-1 in blocked launches of up to per_launch when every
// member allows it
int btstep_batch(const BTMember *d_args, const BTMember *h_args, int n, hipStream_t s, int ty, int per_launch)
{
    MOM6H_REQUIRE(ty == 4 || ty == 8 || ty == 16, "btstep batch: tile rows must be 4, 8 or 16");
    MOM6H_REQUIRE(per_launch >= 1 && per_launch <= 5, "btstep batch: 1..5 substeps per launch");
    const int BT_TY = ty;
    MOM6H_REQUIRE(n >= 1 && n <= kMaxMembers, "btstep batch: 1..%d members", kMaxMembers);
    const int nstep = h_args[0].p.nstep;
    BatchMap mc{}, mt{}, mm{};
    mc.n = mt.n = mm.n = n;
    bool blocked = per_launch > 1, rowwise = true;
    for (int m = 0; m < n; ++m) {
        const Dims &d = h_args[m].p.d;
        MOM6H_REQUIRE(h_args[m].p.nstep == nstep, "btstep batch: members differ in bt_nsteps");
        mc.nbx[m] = (int)((d.s2 + 255) / 256);
        mc.first[m + 1] = mc.first[m] + mc.nbx[m];
        mt.nbx[m] = (d.n1 + BT_TX - 1) / BT_TX;
        mt.nby[m] = (d.n2 + BT_TY - 1) / BT_TY;
        mt.first[m + 1] = mt.first[m] + mt.nbx[m] * mt.nby[m];
        const dim3 gm = bt_multi_grid(d, per_launch);
        mm.nbx[m] = (int)gm.x;
        mm.nby[m] = (int)gm.y;
        mm.first[m + 1] = mm.first[m] + mm.nbx[m] * mm.nby[m];
        blocked = blocked && h_args[m].blocked_ok;
        rowwise = rowwise && h_args[m].m.Datv_row != nullptr;
    }
    k_bt_copy_in_batch<<<(unsigned)mc.first[n], 256, 0, s>>>(d_args, mc);
    MOM6H_LAUNCH_CHECK();
    int from_a = 1;
    for (int q = 1; q <= nstep;) {
        if (blocked && q >= 2 && q <= nstep - 1) {
            const int ns = std::min(per_launch, nstep - q);
            const unsigned nb = (unsigned)mm.first[n], nt = (unsigned)bt_multi_threads(per_launch);
            if (rowwise) {
                if (per_launch == 2) k_bt_multi_batch<2, true><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else if (per_launch == 3) k_bt_multi_batch<3, true><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else if (per_launch == 4) k_bt_multi_batch<4, true><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else k_bt_multi_batch<5, true><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
            } else {
                if (per_launch == 2) k_bt_multi_batch<2, false><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else if (per_launch == 3) k_bt_multi_batch<3, false><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else if (per_launch == 4) k_bt_multi_batch<4, false><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
                else k_bt_multi_batch<5, false><<<nb, nt, 0, s>>>(d_args, mm, q, ns, from_a);
            }
            MOM6H_LAUNCH_CHECK();
            q += ns;
        } else {
            if (ty == 4) k_bt_substep_batch<4><<<(unsigned)mt.first[n], 256, 0, s>>>(d_args, mt, q, from_a);
            else if (ty == 16) k_bt_substep_batch<16><<<(unsigned)mt.first[n], 256, 0, s>>>(d_args, mt, q, from_a);
            else k_bt_substep_batch<8><<<(unsigned)mt.first[n], 256, 0, s>>>(d_args, mt, q, from_a);
            MOM6H_LAUNCH_CHECK();
            ++q;
        }
        from_a = !from_a;
    }
    return MOM6H_OK;
}

// The member record of one btstep call (inputs in, outputs out) on a uniform grid
int bt_member(const mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in, const double *vbt_in,
              double *u_av, double *v_av, double *eta_av, BTMember *out)
{
    MOM6H_REQUIRE(cs->uniform, "btstep batch: members need uniform barotropic coefficients");
    BTMember a{};
    a.p.d = cs->d;
    a.p.nstep = cs->nstep; a.p.first_direction = cs->first_direction;
    a.p.px = cs->halo.px; a.p.py = cs->halo.py;
    a.p.extended = (a.p.px || a.p.py) ? 1 : 0;
    a.p.dtbt = cs->dtbt; a.p.bebt = cs->bebt; a.p.dgeo_de = cs->dgeo_de;
    a.p.ubt_av = cs->ubt_av; a.p.vbt_av = cs->vbt_av;
    a.m = BTUniform{cs->u_Datu, cs->u_gtot, cs->u_rem_u, cs->u_rem_v, cs->u_IareaT, cs->u_IdxCu, cs->u_IdyCv,
                    cs->Datv, cs->f4};
    a.m.Datv_row = cs->use_rows ? cs->Datv_row : nullptr;
    a.m.f4_row = cs->use_rows ? cs->f4_row : nullptr;
    a.blocked_ok = cs->blocked_ok;
    a.A[0] = cs->eta; a.A[1] = cs->ubt; a.A[2] = cs->vbt;
    a.B[0] = cs->eta_b; a.B[1] = cs->ubt_b; a.B[2] = cs->vbt_b;
    a.in[0] = eta_in; a.in[1] = ubt_in; a.in[2] = vbt_in;
    a.out[0] = u_av; a.out[1] = v_av; a.out[2] = eta_av;
    *out = a;
    return MOM6H_OK;
}

}  // namespace mom6h

using namespace mom6h;

extern "C" {

int mom6h_barotropic_init(mom6h_barotropic_cs **out, int isd, int ied, int jsd, int jed, int isc,
                          int iec, int jsc, int jec, int nstep, double dt, double bebt,
                          int first_direction, const double *Datu, const double *Datv,
                          const double *gtot_E, const double *gtot_W, const double *gtot_N,
                          const double *gtot_S, const double *f_4_u, const double *f_4_v,
                          const double *bt_rem_u, const double *bt_rem_v, const double *IareaT,
                          const double *IdxCu, const double *IdyCv, int periodic_x, int periodic_y)
{
    MOM6H_REQUIRE(out != nullptr, "mom6h_barotropic_init: NULL cs");
    MOM6H_REQUIRE(isd < isc && iec < ied && jsd < jsc && jec < jed && isc < iec && jsc < jec,
                  "mom6h_barotropic_init: needs one halo cell on every side");
    MOM6H_REQUIRE(nstep >= 1 && dt > 0.0, "mom6h_barotropic_init: nstep >= 1 and dt > 0 required");
    MOM6H_REQUIRE(!periodic_x || (iec - isc + 1) >= (isc - isd) + (ied - iec),
                  "mom6h_barotropic_init: periodic x needs a compute width of at least both halos");
    MOM6H_REQUIRE(!periodic_y || (jec - jsc + 1) >= (jsc - jsd) + (jed - jec),
                  "mom6h_barotropic_init: periodic y needs a compute height of at least both halos");
    MOM6H_REQUIRE(Datu && Datv && gtot_E && gtot_W && gtot_N && gtot_S && f_4_u && f_4_v && bt_rem_u &&
                      bt_rem_v && IareaT && IdxCu && IdyCv,
                  "mom6h_barotropic_init: NULL array");
    auto *cs = new mom6h_barotropic_cs();
    cs->tile_rows = 4;
    cs->substeps = 5;
    cs->use_rows = 1;
    cs->d = make_dims(isd, ied, jsd, jed, isc, iec, jsc, jec, 1);
    cs->halo = Halo{periodic_x, periodic_y};
    cs->nstep = nstep;
    cs->first_direction = first_direction;
    cs->dtbt = dt / (double)nstep;   // barotropic_init (:72-75)
    cs->bebt = bebt;
    cs->dgeo_de = 1.0;
    const size_t n2 = (size_t)cs->d.s2;
    int rc = 0;
    const double *src[] = {Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv};
    double **dst[] = {&cs->Datu, &cs->Datv, &cs->gtot_E, &cs->gtot_W, &cs->gtot_N, &cs->gtot_S,
                      &cs->bt_rem_u, &cs->bt_rem_v, &cs->IareaT, &cs->IdxCu, &cs->IdyCv};
    for (size_t q = 0; q < sizeof(src) / sizeof(src[0]); ++q) rc = rc ? rc : dev_upload(dst[q], src[q], n2);
    rc = rc ? rc : dev_upload(&cs->f_4_u, f_4_u, 4 * n2);
    rc = rc ? rc : dev_upload(&cs->f_4_v, f_4_v, 4 * n2);
    for (double **pp : {&cs->eta, &cs->ubt, &cs->vbt, &cs->eta_b, &cs->ubt_b, &cs->vbt_b, &cs->ubt_av, &cs->vbt_av})
        rc = rc ? rc : dev_alloc(pp, n2, true);
    // uniform coefficients: every value of a field equal, and the four Coriolis
    // components equal to each other and between u and v points
    if (!rc) {
        auto same = [&](const double *a, double v) {
            for (size_t q = 0; q < n2; ++q)
                if (a[q] != v) return false;
            return true;
        };
        bool uni = same(Datu, Datu[0]) && same(gtot_E, gtot_E[0]) && same(gtot_W, gtot_E[0]) &&
                   same(gtot_N, gtot_E[0]) && same(gtot_S, gtot_E[0]) && same(bt_rem_u, bt_rem_u[0]) &&
                   same(bt_rem_v, bt_rem_v[0]) && same(IareaT, IareaT[0]) && same(IdxCu, IdxCu[0]) &&
                   same(IdyCv, IdyCv[0]);
        std::vector<double> f(n2);
        for (size_t q = 0; uni && q < n2; ++q) {
            f[q] = f_4_u[4 * q];
            for (int c = 0; c < 4; ++c)
                if (f_4_u[4 * q + c] != f[q] || f_4_v[4 * q + c] != f[q]) uni = false;
        }
        if (uni) {
            // the blocked kernel reads Datv of a periodic halo row at its image
            // (the row the reference refills it from): it runs only when they agree
            bool images = true;
            if (periodic_y) {
                const int n1 = ied - isd + 1;
                for (int i = isc; images && i <= iec; ++i)
                    for (int j = jsd; images && j < jsc; ++j)
                        images = Datv[(size_t)(j - jsd) * n1 + (i - isd)] ==
                                 Datv[(size_t)(j + (jec - jsc + 1) - jsd) * n1 + (i - isd)];
            }
            cs->blocked_ok = images ? 1 : 0;
            // Datv and f4 constant along every row of the compute domain: the
            // blocked kernel stages one value per row
            {
                const int n1 = ied - isd + 1, n2r = jed - jsd + 1;
                bool rows = true;
                std::vector<double> dvr(n2r), f4r(n2r);
                for (int j = 0; j < n2r && rows; ++j) {
                    const size_t r0 = (size_t)j * n1;
                    dvr[j] = Datv[r0 + (isc - isd)];
                    f4r[j] = f[r0 + (isc - isd)];
                    for (int i = isc; i <= iec && rows; ++i)
                        rows = Datv[r0 + (i - isd)] == dvr[j] && f[r0 + (i - isd)] == f4r[j];
                }
                if (rows) {
                    rc = dev_upload(&cs->Datv_row, dvr.data(), (size_t)n2r);
                    rc = rc ? rc : dev_upload(&cs->f4_row, f4r.data(), (size_t)n2r);
                }
            }
            cs->uniform = 1;
            cs->u_Datu = Datu[0]; cs->u_gtot = gtot_E[0]; cs->u_rem_u = bt_rem_u[0]; cs->u_rem_v = bt_rem_v[0];
            cs->u_IareaT = IareaT[0]; cs->u_IdxCu = IdxCu[0]; cs->u_IdyCv = IdyCv[0];
            rc = rc ? rc : dev_upload(&cs->f4, f.data(), n2);
        }
    }
    if (rc) {
        mom6h_barotropic_end(cs);
        return rc;
    }
    *out = cs;
    return MOM6H_OK;
}

}  // extern "C"

namespace mom6h {

// Scratch of mom6h_btstep_noalloc: the two state sets and the two velocity
// accumulators, 8 planes of n1*n2 values packed from the start, for grids of
// at most kNoallocPoints points (n1*n2).  Each GPU holds one copy for the
// life of the process.  Host code must not pass the symbol itself to a
// kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress gives for the current device.
constexpr long long kNoallocPoints = 1LL << 20;
static __device__ __align__(256) double scratch[8 * kNoallocPoints];

// device address of `scratch` on the current device, looked up once per device
static int noalloc_scratch(double **out)
{
    static std::mutex mu;
    static double *ptr[64] = {};
    int dev = 0;
    MOM6H_CHECK(hipGetDevice(&dev));
    MOM6H_REQUIRE(dev >= 0 && dev < 64, "mom6h_btstep_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        MOM6H_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MOM6H_REQUIRE(p != nullptr, "hipGetSymbolAddress(scratch) gave no address");
        ptr[dev] = (double *)p;
    }
    *out = ptr[dev];
    return MOM6H_OK;
}

// One btstep call with its state sets A = (eta, ubt, vbt), B = (eta_b, ubt_b,
// vbt_b) and accumulators (ubt_av, vbt_av) in the 8 planes w[0..7]; zero_b:
// clear B on the way in (scratch whose previous contents belong to another call)
static int btstep_run(mom6h_barotropic_cs *cs, double *const w[8], bool zero_b, const double *eta_in,
                      const double *ubt_in, const double *vbt_in, double *u_av, double *v_av, double *eta_av,
                      hipStream_t strm)
{
    BT p;
    p.d = cs->d;
    p.nstep = cs->nstep; p.first_direction = cs->first_direction;
    p.px = cs->halo.px; p.py = cs->halo.py;
    p.extended = (p.px || p.py) ? 1 : 0;
    p.dtbt = cs->dtbt; p.bebt = cs->bebt; p.dgeo_de = cs->dgeo_de;
    p.ubt_av = w[6]; p.vbt_av = w[7];
    // copy-in, then one launch per substep ping-ponging between the two state sets
    const Dims &d = cs->d;
    {
        const int bs = 256;
        k_bt_copy_in<<<(unsigned)((d.s2 + bs - 1) / bs), bs, 0, strm>>>(
            d, eta_in, ubt_in, vbt_in, w[0], w[1], w[2], w[6], w[7], zero_b ? w[3] : nullptr,
            zero_b ? w[4] : nullptr, zero_b ? w[5] : nullptr);
        MOM6H_LAUNCH_CHECK();
    }
    const int BT_TY = cs->tile_rows;
    dim3 grid((d.n1 + BT_TX - 1) / BT_TX, (d.n2 + BT_TY - 1) / BT_TY), block(256);
    const dim3 grid_m = bt_multi_grid(d, cs->substeps);
    const int per_launch = cs->substeps;
    double *A[3] = {w[0], w[1], w[2]}, *B[3] = {w[3], w[4], w[5]};
    for (int n = 1; n <= cs->nstep;) {
        // substeps 2 .. nstep-1 in blocks of up to per_launch
        if (per_launch > 1 && cs->blocked_ok && n >= 2 && n <= cs->nstep - 1) {
            BTMulti st;
            st.eta_o = A[0]; st.ubt_o = A[1]; st.vbt_o = A[2];
            st.eta_n = B[0]; st.ubt_n = B[1]; st.vbt_n = B[2];
            st.n0 = n;
            st.ns = std::min(per_launch, cs->nstep - n);
            BTUniform m{cs->u_Datu, cs->u_gtot, cs->u_rem_u, cs->u_rem_v, cs->u_IareaT, cs->u_IdxCu,
                        cs->u_IdyCv, cs->Datv, cs->f4};
            if (cs->Datv_row && cs->use_rows) {
                m.Datv_row = cs->Datv_row;
                m.f4_row = cs->f4_row;
                launch_bt_multi<true>(per_launch, grid_m, p, m, st, strm);
            } else {
                launch_bt_multi<false>(per_launch, grid_m, p, m, st, strm);
            }
            MOM6H_LAUNCH_CHECK();
            for (int q = 0; q < 3; ++q) std::swap(A[q], B[q]);
            n += st.ns;
            continue;
        }
        BTStep st;
        st.eta_o = A[0]; st.ubt_o = A[1]; st.vbt_o = A[2];
        st.eta_n = B[0]; st.ubt_n = B[1]; st.vbt_n = B[2];
        st.out_u = u_av; st.out_v = v_av; st.out_eta = eta_av;
        st.n = n;
        if (cs->uniform) {
            BTUniform m{cs->u_Datu, cs->u_gtot, cs->u_rem_u, cs->u_rem_v, cs->u_IareaT, cs->u_IdxCu, cs->u_IdyCv,
                        cs->Datv, cs->f4};
            if (BT_TY == 4) k_bt_substep<BTUniform, 4><<<grid, block, 0, strm>>>(p, m, st);
            else if (BT_TY == 16) k_bt_substep<BTUniform, 16><<<grid, block, 0, strm>>>(p, m, st);
            else k_bt_substep<BTUniform, 8><<<grid, block, 0, strm>>>(p, m, st);
        } else {
            BTArrays m{cs->Datu, cs->Datv, cs->gtot_E, cs->gtot_W, cs->gtot_N, cs->gtot_S, cs->f_4_u, cs->f_4_v,
                       cs->bt_rem_u, cs->bt_rem_v, cs->IareaT, cs->IdxCu, cs->IdyCv};
            if (BT_TY == 4) k_bt_substep<BTArrays, 4><<<grid, block, 0, strm>>>(p, m, st);
            else if (BT_TY == 16) k_bt_substep<BTArrays, 16><<<grid, block, 0, strm>>>(p, m, st);
            else k_bt_substep<BTArrays, 8><<<grid, block, 0, strm>>>(p, m, st);
        }
        MOM6H_LAUNCH_CHECK();
        for (int q = 0; q < 3; ++q) std::swap(A[q], B[q]);
        ++n;
    }
    return MOM6H_OK;
}

}  // namespace mom6h

extern "C" {

int mom6h_btstep(mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in,
                 const double *vbt_in, double *u_av, double *v_av, double *eta_av, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_btstep: NULL cs");
    MOM6H_REQUIRE(eta_in && ubt_in && vbt_in && u_av && v_av && eta_av, "mom6h_btstep: NULL array");
    double *const w[8] = {cs->eta, cs->ubt, cs->vbt, cs->eta_b, cs->ubt_b, cs->vbt_b, cs->ubt_av, cs->vbt_av};
    return btstep_run(cs, w, false, eta_in, ubt_in, vbt_in, u_av, v_av, eta_av, (hipStream_t)stream);
}

int mom6h_btstep_noalloc(mom6h_barotropic_cs *cs, const double *eta_in, const double *ubt_in,
                         const double *vbt_in, double *u_av, double *v_av, double *eta_av, void *stream)
{
    MOM6H_REQUIRE(cs != nullptr, "mom6h_btstep_noalloc: NULL cs");
    MOM6H_REQUIRE(eta_in && ubt_in && vbt_in && u_av && v_av && eta_av, "mom6h_btstep_noalloc: NULL array");
    if (cs->d.s2 > kNoallocPoints)
        return set_error(MOM6H_EINVAL,
                         "mom6h_btstep_noalloc: the grid has %lld points (n1*n2), the scratch holds at most %lld",
                         cs->d.s2, kNoallocPoints);
    double *base = nullptr;
    const int rc = noalloc_scratch(&base);
    if (rc) return rc;
    const long long n = cs->d.s2;
    double *const w[8] = {base, base + n, base + 2 * n, base + 3 * n, base + 4 * n, base + 5 * n, base + 6 * n,
                          base + 7 * n};
    return btstep_run(cs, w, true, eta_in, ubt_in, vbt_in, u_av, v_av, eta_av, (hipStream_t)stream);
}

int mom6h_barotropic_end(mom6h_barotropic_cs *cs)
{
    if (!cs) return MOM6H_OK;
    for (double *p : {cs->Datu, cs->Datv, cs->gtot_E, cs->gtot_W, cs->gtot_N, cs->gtot_S, cs->f_4_u,
                      cs->f_4_v, cs->bt_rem_u, cs->bt_rem_v, cs->IareaT, cs->IdxCu, cs->IdyCv, cs->f4,
                      cs->Datv_row, cs->f4_row,
                      cs->eta, cs->ubt, cs->vbt, cs->eta_b, cs->ubt_b, cs->vbt_b, cs->ubt_av, cs->vbt_av})
        dev_free(p);
    delete cs;
    return MOM6H_OK;
}

}  // extern "C"
