// mom6h_fused.hip -- fused momentum stencil for grids with uniform metrics.
//
// One launch performs, for every layer, what the reference does in four
// module calls (app/rk2_serial_driver.F90:236-270 and :316-350):
//   compute_transports (:544-561) -> hor_visc (src/serial/mom6_hor_visc_serial.F90,
//   Laplacian) -> CorAdCalc (src/serial/mom6_coriolis_serial.F90, Sadourny
//   energy scheme) -> velocity update  ou = base + fdt*(CA + diff)
// The transports, tendencies and potential vorticity never leave the chip: a
// block stages u, v, h of its tile plus a one-cell ring in LDS for one layer,
// forms q, KE and the viscous stresses on the staged q/h points (each once),
// then every thread finishes its velocity points.  It walks all layers, so the 2D masks and Coriolis
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

// Output tile MT_X x TY (TY = 4, 8 or 16 rows, chosen at launch); the staged
// region is [i0-1, i1+1] x [j0-1, j1+1]
constexpr int MT_X = 64;
constexpr int MW = MT_X + 2;

// Biharmonic viscosity, first half: Del2u on u points I = is-1..ie,
// j = js..je and Del2v on v points i = is..ie, J = js-1..je
// (mom6_hor_visc_serial.F90:587-607, uniform metrics).  The arrays are zero
// at allocation and never written outside those ranges, which is what the
// reference's second half reads there.
template <typename T>
__device__ __forceinline__ void del2_body(const MomArgsT<T> &a, int bx, int by, int bz)
{
    const Dims &d = a.d;
    const UMetT<T> g(a.g);
    const int i = d.isc - 1 + bx * 64 + (threadIdx.x & 63);
    const int j = d.jsc - 1 + by * 4 + (threadIdx.x >> 6);
    const int k = 1 + bz;
    if (i > d.iec || j > d.jec) return;
    auto U = [&](int ii, int jj) -> T { return a.u[IX3(d, ii, jj, k)]; };
    auto V = [&](int ii, int jj) -> T { return a.v[IX3(d, ii, jj, k)]; };
    auto sh_xx = [&](int ii, int jj) -> T {   // h point (:503-510)
        const T dudx = g.DY_dxT * ((g.IdyCu * U(ii, jj)) - (g.IdyCu * U(ii - 1, jj)));
        const T dvdy = g.DX_dyT * ((g.IdxCv * V(ii, jj)) - (g.IdxCv * V(ii, jj - 1)));
        return dudx - dvdy;
    };
    auto sh_xy = [&](int I, int J) -> T {    // q point (:513-547)
        const T dvdx = g.DY_dxBu * ((V(I + 1, J) * g.IdyCv) - (V(I, J) * g.IdyCv));
        const T dudy = g.DX_dyBu * ((U(I, J + 1) * g.IdxCu) - (U(I, J) * g.IdxCu));
        const T m = (T)a.mask2dBu[IX2(d, I, J)];
        return (a.no_slip ? (T(2.0) - m) : m) * (dvdx + dudy);
    };
    const long long c = IX3(d, i, j, k);
    if (j >= d.jsc)
        a.del2u[c] = g.Idx2dyCu * ((g.dx2q * sh_xy(i, j)) - (g.dx2q * sh_xy(i, j - 1))) +
                     g.Idxdy2u * ((g.dy2h * sh_xx(i + 1, j)) - (g.dy2h * sh_xx(i, j)));
    if (i >= d.isc)
        a.del2v[c] = g.Idxdy2v * ((g.dy2q * sh_xy(i, j)) - (g.dy2q * sh_xy(i - 1, j))) -
                     g.Idx2dyCv * ((g.dx2h * sh_xx(i, j + 1)) - (g.dx2h * sh_xx(i, j)));
}

template <typename T>
__global__ void __launch_bounds__(256) k_del2(MomArgsT<T> a)
{
    del2_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <typename T>
__global__ void __launch_bounds__(256) k_del2_batch(const MomArgsT<T> *__restrict__ args, BatchMap mp)
{
    int m, local;
    batch_locate(mp, blockIdx.x, m, local);
    const int r = local / mp.nbx[m];
    const MomArgsT<T> a = args[m];
    del2_body(a, local % mp.nbx[m], r % mp.nby[m], r / mp.nby[m]);
}

// Division in the stencil's type: exact in FP64 (the reference's arithmetic);
// in the FP32 form a product with the hardware reciprocal (about 1 ulp, the
// rounding scale of the FP32 stencil itself)
template <typename T>
__device__ __forceinline__ T sdiv(T num, T den)
{
    if constexpr (sizeof(T) == 4) return num * __builtin_amdgcn_rcpf(den);
    else return num / den;
}

// The tile (tile_x, tile_y) of one grid.  BIH: the biharmonic stress of the
// Del2 fields staged like u and v, added to the Laplacian stress
// (:824-886, :979-1036).
template <typename T, bool BIH, int TY>
__device__ __forceinline__ void momentum_body(const MomArgsT<T> &a, int tile_x, int tile_y, int k_first, int k_last)
{
    constexpr int MH = TY + 2, RPT = TY / 4;   // staged rows; output rows per thread
    __shared__ T s_u[MH][MW], s_v[MH][MW], s_h[MH][MW];
    __shared__ T s_q[MH][MW], s_KE[MH][MW];
    // the Laplacian (+ biharmonic) stresses str_xx*h at the staged h points and
    // str_xy*hq at the staged q points, formed once per point (each is read by
    // up to three output points)
    __shared__ T s_sxx[MH][MW], s_sxy[MH][MW];
    __shared__ T s_d2u[BIH ? MH : 1][BIH ? MW : 1], s_d2v[BIH ? MH : 1][BIH ? MW : 1];
    // the 2D masks and Coriolis of the staged region, the same for every
    // layer: the masks as bytes (the host checks that they hold 0 or 1) and
    // Coriolis as one value per row (the host checks that CoriolisBu is
    // constant along every row), which leaves a 64 x 8 tile in FP64 38 KB of
    // LDS, four blocks per CU instead of three
    __shared__ unsigned char s_mT[MH][MW], s_mBu[MH][MW];
    __shared__ T s_cor[MH];
    const Dims &d = a.d;
    const UMetT<T> g(a.g);
    const int is = d.isc, ie = d.iec, js = d.jsc, je = d.jec;
    const int i0 = is + tile_x * MT_X, j0 = js + tile_y * TY;
    const int bi = i0 - 1, bj = j0 - 1;
    const int tid = threadIdx.x;
    const int tx = tid % MT_X, ty = tid / MT_X;          // 64 x 4 threads, RPT rows each

    // u, v, h of the staged region [i0-1, i1+1] x [j0-1, j1+1] of one layer go
    // through registers: the next layer's loads are issued before the current
    // layer is computed (NS points per thread), and so are the base velocities
    // of the thread's output rows
    constexpr int NS = (MW * MH + 255) / 256;
    T pu[NS], pv[NS], ph[NS], pbu[RPT], pbv[RPT];
    T p2u[BIH ? NS : 1], p2v[BIH ? NS : 1];
    auto load_layer = [&](int k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            const int li = t % MW, lj = t / MW, i = bi + li, j = bj + lj;
            pu[r] = T(0.0); pv[r] = T(0.0); ph[r] = T(0.0);
            if (BIH) { p2u[BIH ? r : 0] = T(0.0); p2v[BIH ? r : 0] = T(0.0); }
            if (t < MW * MH && i <= d.ied && j <= d.jed) {
                const long long c = IX3(d, i, j, k);
                pu[r] = a.u[c]; pv[r] = a.v[c]; ph[r] = a.h[c];
                if (BIH) { p2u[BIH ? r : 0] = a.del2u[c]; p2v[BIH ? r : 0] = a.del2v[c]; }
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            pbu[r] = T(0.0); pbv[r] = T(0.0);
            if (i <= ie && j <= je) {
                const long long c = IX3(d, i, j, k);
                pbu[r] = a.bu[c]; pbv[r] = a.bv[c];
            }
        }
    };
    for (int t = tid; t < MW * MH; t += blockDim.x) {
        const int li = t % MW, lj = t / MW, i = bi + li, j = bj + lj;
        T mt = T(0.0), mb = T(0.0), cor = T(0.0);
        if (i <= d.ied && j <= d.jed) {
            const long long c2 = IX2(d, i, j);
            mt = (T)a.mask2dT[c2];
            mb = (T)a.mask2dBu[c2];
            cor = (T)a.CoriolisBu[c2];
        }
        s_mT[lj][li] = (unsigned char)(mt != T(0.0));
        s_mBu[lj][li] = (unsigned char)(mb != T(0.0));
        if (li == 0) s_cor[lj] = cor;   // i = i0-1: inside the data domain
    }
    // Laplacian stress pieces (mom6_hor_visc_serial.F90:503-547, 807-1053)
    auto str_xx = [&](int pi, int pj) -> T {   // LDS coords of an h point
        const T dudx = g.DY_dxT * ((g.IdyCu * s_u[pj][pi]) - (g.IdyCu * s_u[pj][pi - 1]));
        const T dvdy = g.DX_dyT * ((g.IdxCv * s_v[pj][pi]) - (g.IdxCv * s_v[pj - 1][pi]));
        T str = -(a.Kh * (dudx - dvdy));
        if (BIH) {
            const int qi = BIH ? pi : 0, qj = BIH ? pj : 0;
            const T d_del2u = (g.IdyCu * s_d2u[qj][qi]) - (g.IdyCu * s_d2u[qj][BIH ? qi - 1 : 0]);
            const T d_del2v = (g.IdxCv * s_d2v[qj][qi]) - (g.IdxCv * s_d2v[BIH ? qj - 1 : 0][qi]);
            str = str + a.Ah * ((g.DY_dxT * d_del2u) - (g.DX_dyT * d_del2v));
        }
        return str * (s_h[pj][pi] * T(1.0));
    };
    auto str_xy = [&](int pi, int pj) -> T {   // q point at LDS (pi, pj)
        const T dvdx = g.DY_dxBu * ((s_v[pj][pi + 1] * g.IdyCv) - (s_v[pj][pi] * g.IdyCv));
        const T dudy = g.DX_dyBu * ((s_u[pj + 1][pi] * g.IdxCu) - (s_u[pj][pi] * g.IdxCu));
        const T m = (T)s_mBu[pj][pi];
        const T sh = (a.no_slip ? (T(2.0) - m) : m) * (dvdx + dudy);
        const T hq = T(0.25) * ((s_h[pj][pi] + s_h[pj + 1][pi + 1]) + (s_h[pj][pi + 1] + s_h[pj + 1][pi]));
        T str = -(a.Kh * sh);
        if (BIH) {
            const int qi = BIH ? pi : 0, qj = BIH ? pj : 0;
            str = str + a.Ah * ((g.DY_dxBu * ((s_d2v[qj][BIH ? qi + 1 : 0] * g.IdyCv) - (s_d2v[qj][qi] * g.IdyCv))) +
                                (g.DX_dyBu * ((s_d2u[BIH ? qj + 1 : 0][qi] * g.IdxCu) - (s_d2u[qj][qi] * g.IdxCu))));
        }
        return a.no_slip ? str * (hq * T(1.0)) : str * (hq * m * T(1.0));
    };
    load_layer(k_first);
    for (int k = k_first; k <= k_last; ++k) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int t = tid + 256 * r;
            if (t < MW * MH) {
                const int li = t % MW, lj = t / MW;
                s_u[lj][li] = pu[r]; s_v[lj][li] = pv[r]; s_h[lj][li] = ph[r];
                if (BIH) { s_d2u[BIH ? lj : 0][BIH ? li : 0] = p2u[BIH ? r : 0]; s_d2v[BIH ? lj : 0][BIH ? li : 0] = p2v[BIH ? r : 0]; }
            }
        }
        T bu_r[RPT], bv_r[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) { bu_r[r] = pbu[r]; bv_r[r] = pbv[r]; }
        __syncthreads();
        if (k < k_last) load_layer(k + 1);
        // q on q points [i0-1, i1] x [j0-1, j1] (mom6_coriolis_serial.F90:134-168)
        // and KE on h points [i0, i1+1] x [j0, j1+1] (:196-203)
        for (int t = tid; t < MW * MH; t += blockDim.x) {
            const int li = t % MW, lj = t / MW, i = bi + li, j = bj + lj;
            if (li < MW - 1 && lj < MH - 1 && i <= ie && j <= je) {
                const T dvdx = (s_v[lj][li + 1] * g.dyCv) - (s_v[lj][li] * g.dyCv);
                const T dudy = (s_u[lj + 1][li] * g.dxCu) - (s_u[lj][li] * g.dxCu);
                const T hAu0 = (j < js) ? T(0.0) : T(0.5) * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj][li + 1]));
                const T hAu1 = T(0.5) * ((g.areaT * s_h[lj + 1][li]) + (g.areaT * s_h[lj + 1][li + 1]));
                const T hAv0 = (i < is) ? T(0.0) : T(0.5) * ((g.areaT * s_h[lj][li]) + (g.areaT * s_h[lj + 1][li]));
                const T hAv1 = T(0.5) * ((g.areaT * s_h[lj][li + 1]) + (g.areaT * s_h[lj + 1][li + 1]));
                const T rel_vort = (T)s_mBu[lj][li] * (dvdx - dudy) * g.IareaBu;
                const T abs_vort = s_cor[lj] + rel_vort;
                const T hArea_q = (hAu0 + hAu1) + (hAv0 + hAv1);
                s_q[lj][li] = abs_vort * sdiv(g.Area_q, hArea_q + T(1.0e-20));
            }
            if (li >= 1 && lj >= 1) s_sxx[lj][li] = str_xx(li, lj);
            if (li < MW - 1 && lj < MH - 1) s_sxy[lj][li] = str_xy(li, lj);
            if (li >= 1 && lj >= 1 && i <= ie && j <= je) {
                const T uc = s_u[lj][li], uw = s_u[lj][li - 1], vc = s_v[lj][li], vs = s_v[lj - 1][li];
                s_KE[lj][li] = sdiv(T(0.25) * ((g.dyCu * (uc * uc) + g.dyCu * (uw * uw)) +
                                            (g.dxCv * (vc * vc) + g.dxCv * (vs * vs))), g.areaT);
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int r = 0; r < RPT; ++r) {
            const int i = i0 + tx, j = j0 + ty + 4 * r;
            if (i > ie || j > je) continue;
            const int li = i - bi, lj = j - bj;
            const long long c = IX3(d, i, j, k);
            if (i <= ie - 1) {
                // diffu (:1059-1067)
                const T h_u = T(0.5) * ((T)s_mT[lj][li] * s_h[lj][li] + (T)s_mT[lj][li + 1] * s_h[lj][li + 1]);
                const T diffu = sdiv((g.IdxCu * ((g.dx2q * s_sxy[lj - 1][li]) - (g.dx2q * s_sxy[lj][li])) +
                                      g.IdyCu * ((g.dy2h * s_sxx[lj][li]) - (g.dy2h * s_sxx[lj][li + 1]))) *
                                         g.IareaCu, h_u + a.h_neglect);
                // CAu, Sadourny (:208-216); vh from compute_transports
                auto vh = [&](int pi, int pj) { return s_v[pj][pi] * T(0.5) * (s_h[pj][pi] + s_h[pj + 1][pi]) * g.dxCv; };
                const T KEx = (s_KE[lj][li + 1] - s_KE[lj][li]) * g.IdxCu;
                const T CAu = T(0.25) * ((s_q[lj][li] * (vh(li + 1, lj) + vh(li, lj))) +
                                           (s_q[lj - 1][li] * (vh(li, lj - 1) + vh(li + 1, lj - 1)))) * g.IdxCu - KEx;
                a.ou[c] = bu_r[r] + a.fdt * (CAu + diffu);
            }
            if (j <= je - 1) {
                // diffv (:1070-1078)
                const T h_v = T(0.5) * ((T)s_mT[lj][li] * s_h[lj][li] + (T)s_mT[lj + 1][li] * s_h[lj + 1][li]);
                const T diffv = sdiv((g.IdyCv * ((g.dy2q * s_sxy[lj][li - 1]) - (g.dy2q * s_sxy[lj][li])) -
                                      g.IdxCv * ((g.dx2h * s_sxx[lj][li]) - (g.dx2h * s_sxx[lj + 1][li]))) *
                                         g.IareaCv, h_v + a.h_neglect);
                // CAv, Sadourny (:218-226); uh from compute_transports
                auto uh = [&](int pi, int pj) { return s_u[pj][pi] * T(0.5) * (s_h[pj][pi] + s_h[pj][pi + 1]) * g.dyCu; };
                const T KEy = (s_KE[lj + 1][li] - s_KE[lj][li]) * g.IdyCv;
                const T CAv = -(T(0.25) * ((s_q[lj][li - 1] * (uh(li - 1, lj) + uh(li - 1, lj + 1))) +
                                             (s_q[lj][li] * (uh(li, lj) + uh(li, lj + 1)))) * g.IdyCv) - KEy;
                a.ov[c] = bv_r[r] + a.fdt * (CAv + diffv);
            }
        }
        __syncthreads();
    }
}

// layers k_first..k_last of split kz of nkz (the layers are independent)
__device__ inline void layer_range(int nk, int kz, int nkz, int &k_first, int &k_last)
{
    const int per = (nk + nkz - 1) / nkz;
    k_first = 1 + kz * per;
    k_last = k_first + per - 1 < nk ? k_first + per - 1 : nk;
}

// grid (tiles x, tiles y, layer splits); each block walks its split's layers
template <typename T, bool BIH, int TY>
__global__ void __launch_bounds__(256) k_momentum(MomArgsT<T> a)
{
    const int nxy = gridDim.x * gridDim.y;
    const int tile = xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), nxy * gridDim.z);
    const int xy = tile % nxy;
    int k0, k1;
    layer_range(a.d.nk, tile / nxy, gridDim.z, k0, k1);
    if (k0 > k1) return;
    momentum_body<T, BIH, TY>(a, xy % gridDim.x, xy / gridDim.x, k0, k1);
}

// every member's tiles in one launch, renumbered per XCD over the whole batch
template <typename T, bool BIH, int TY>
__global__ void __launch_bounds__(256) k_momentum_batch(const MomArgsT<T> *__restrict__ args, BatchMap mp)
{
    int m, local;
    batch_locate(mp, xcd_tile(blockIdx.x, gridDim.x), m, local);
    const MomArgsT<T> a = args[m];
    const int nxy = mp.nbx[m] * mp.nby[m], xy = local % nxy;
    int k0, k1;
    layer_range(a.d.nk, local / nxy, a.nkz, k0, k1);
    if (k0 > k1) return;
    momentum_body<T, BIH, TY>(a, xy % mp.nbx[m], xy / mp.nbx[m], k0, k1);
}

template <typename T, bool BIH>
void momentum_launch_ty(const MomArgsT<T> &a, int ty, dim3 grid, hipStream_t s)
{
    if (ty == 4) k_momentum<T, BIH, 4><<<grid, 256, 0, s>>>(a);
    else if (ty == 16) k_momentum<T, BIH, 16><<<grid, 256, 0, s>>>(a);
    else k_momentum<T, BIH, 8><<<grid, 256, 0, s>>>(a);
}

template <typename T>
int launch_momentum(const MomArgsT<T> &a, hipStream_t s, int ty)
{
    const Dims &d = a.d;
    MOM6H_REQUIRE(ty == 4 || ty == 8 || ty == 16, "momentum: tile height %d not in {4, 8, 16}", ty);
    if (a.del2u) {
        dim3 g2((d.iec - d.isc + 2 + 63) / 64, (d.jec - d.jsc + 2 + 3) / 4, d.nk);
        k_del2<T><<<g2, dim3(256), 0, s>>>(a);
        MOM6H_LAUNCH_CHECK();
    }
    const int nkz = a.nkz < 1 ? 1 : (a.nkz > d.nk ? d.nk : a.nkz);
    dim3 grid((d.iec - d.isc + 1 + MT_X - 1) / MT_X, (d.jec - d.jsc + 1 + ty - 1) / ty, nkz);
    if (a.del2u) momentum_launch_ty<T, true>(a, ty, grid, s);
    else momentum_launch_ty<T, false>(a, ty, grid, s);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

template <typename T, bool BIH>
void momentum_batch_ty(const MomArgsT<T> *d_args, const BatchMap &mp, int ty, hipStream_t s)
{
    const unsigned nb = (unsigned)mp.first[mp.n];
    if (ty == 4) k_momentum_batch<T, BIH, 4><<<nb, 256, 0, s>>>(d_args, mp);
    else if (ty == 16) k_momentum_batch<T, BIH, 16><<<nb, 256, 0, s>>>(d_args, mp);
    else k_momentum_batch<T, BIH, 8><<<nb, 256, 0, s>>>(d_args, mp);
}

template <typename T>
int launch_momentum_batch(const MomArgsT<T> *d_args, const MomArgsT<T> *h_args, int n, hipStream_t s, int ty)
{
    MOM6H_REQUIRE(n >= 1 && n <= kMaxMembers, "momentum batch: 1..%d members", kMaxMembers);
    MOM6H_REQUIRE(ty == 4 || ty == 8 || ty == 16, "momentum: tile height %d not in {4, 8, 16}", ty);
    const bool bih = h_args[0].del2u != nullptr;
    for (int m = 1; m < n; ++m)
        MOM6H_REQUIRE((h_args[m].del2u != nullptr) == bih, "momentum batch: members differ in biharmonic");
    BatchMap mp{};
    mp.n = n;
    if (bih) {
        for (int m = 0; m < n; ++m) {
            const Dims &d = h_args[m].d;
            mp.nbx[m] = (d.iec - d.isc + 2 + 63) / 64;
            mp.nby[m] = (d.jec - d.jsc + 2 + 3) / 4;
            mp.first[m + 1] = mp.first[m] + mp.nbx[m] * mp.nby[m] * d.nk;
        }
        k_del2_batch<T><<<(unsigned)mp.first[n], 256, 0, s>>>(d_args, mp);
        MOM6H_LAUNCH_CHECK();
    }
    for (int m = 0; m < n; ++m) {
        const Dims &d = h_args[m].d;
        MOM6H_REQUIRE(h_args[m].nkz >= 1 && h_args[m].nkz <= d.nk, "momentum batch: layer splits 1..nk");
        mp.nbx[m] = (d.iec - d.isc + 1 + MT_X - 1) / MT_X;
        mp.nby[m] = (d.jec - d.jsc + 1 + ty - 1) / ty;
        mp.first[m + 1] = mp.first[m] + mp.nbx[m] * mp.nby[m] * h_args[m].nkz;
    }
    if (bih) momentum_batch_ty<T, true>(d_args, mp, ty, s);
    else momentum_batch_ty<T, false>(d_args, mp, ty, s);
    MOM6H_LAUNCH_CHECK();
    return MOM6H_OK;
}

template int launch_momentum<double>(const MomArgsT<double> &, hipStream_t, int);
template int launch_momentum<float>(const MomArgsT<float> &, hipStream_t, int);
template int launch_momentum_batch<double>(const MomArgsT<double> *, const MomArgsT<double> *, int, hipStream_t, int);
template int launch_momentum_batch<float>(const MomArgsT<float> *, const MomArgsT<float> *, int, hipStream_t, int);

}  // namespace mom6h
