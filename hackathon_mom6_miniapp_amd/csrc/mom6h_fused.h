// mom6h_fused.h -- kernels of the fused step for grids with uniform metrics.
#pragma once

#include "mom6h_internal.h"

namespace mom6h {

// One value per metric array (the arrays of a uniform grid are constant; the
// values are read from the arrays, so every expression sees the same numbers
// as the general kernels do).
struct UMetrics {
    double areaT, IareaT, dxT, IdxT, dyCu, dy_Cu, dxCu, IdxCu, IdyCu, dxCv, dyCv, IdxCv, IdyCv;
    double IareaBu, IareaCu, IareaCv, Area_q;
    double DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, dx2h, dy2h, dx2q, dy2q;
    double Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v;   // biharmonic (hor_visc_init :262-283)
};

// The metric values in the stencil's arithmetic type (T = double: unchanged)
template <typename T>
struct UMetT {
    T areaT, IareaT, dxT, IdxT, dyCu, dy_Cu, dxCu, IdxCu, IdyCu, dxCv, dyCv, IdxCv, IdyCv;
    T IareaBu, IareaCu, IareaCv, Area_q;
    T DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, dx2h, dy2h, dx2q, dy2q;
    T Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v;
    __host__ __device__ explicit UMetT(const UMetrics &g)
        : areaT((T)g.areaT), IareaT((T)g.IareaT), dxT((T)g.dxT), IdxT((T)g.IdxT), dyCu((T)g.dyCu),
          dy_Cu((T)g.dy_Cu), dxCu((T)g.dxCu), IdxCu((T)g.IdxCu), IdyCu((T)g.IdyCu), dxCv((T)g.dxCv),
          dyCv((T)g.dyCv), IdxCv((T)g.IdxCv), IdyCv((T)g.IdyCv), IareaBu((T)g.IareaBu), IareaCu((T)g.IareaCu),
          IareaCv((T)g.IareaCv), Area_q((T)g.Area_q), DY_dxT((T)g.DY_dxT), DX_dyT((T)g.DX_dyT),
          DY_dxBu((T)g.DY_dxBu), DX_dyBu((T)g.DX_dyBu), dx2h((T)g.dx2h), dy2h((T)g.dy2h), dx2q((T)g.dx2q),
          dy2q((T)g.dy2q), Idx2dyCu((T)g.Idx2dyCu), Idxdy2u((T)g.Idxdy2u), Idx2dyCv((T)g.Idx2dyCv),
          Idxdy2v((T)g.Idxdy2v)
    {
    }
};

// Arguments of the fused momentum stencil.  T is the type of the 3D layer
// fields and of the stencil arithmetic (double: the reference's FP64; float:
// the mixed-precision form); 2D masks and Coriolis stay FP64 in memory.
template <typename T>
struct MomArgsT {
    Dims d;
    UMetrics g;
    T Kh, h_neglect, fdt;
    int no_slip;
    const T *u, *v, *h;               // stencil inputs
    const T *bu, *bv;                 // bases of the update (may alias ou/ov)
    T *ou, *ov;
    const double *mask2dT, *mask2dBu, *CoriolisBu;
    // biharmonic viscosity: Ah and the Del2 scratch (NULL: Laplacian only)
    T Ah;
    T *del2u, *del2v;
    int nkz;   // the layers split over this many blocks per tile (1..nk)
};
using MomArgs = MomArgsT<double>;

// One grid (tile height ty = 4, 8 or 16 rows), or every member of an ensemble
// in one launch per kernel (d_args: the members' arguments on the device,
// h_args: the same on the host)
template <typename T>
int launch_momentum(const MomArgsT<T> &a, hipStream_t s, int ty = 8);
template <typename T>
int launch_momentum_batch(const MomArgsT<T> *d_args, const MomArgsT<T> *h_args, int n, hipStream_t s, int ty = 8);

// true and the common value when every element of a[0..n) is identical
inline bool uniform_value(const double *a, size_t n, double *v)
{
    for (size_t q = 1; q < n; ++q)
        if (!(a[q] == a[0])) return false;
    *v = a[0];
    return true;
}

// true when every element of a[0..n) is 0 or 1 (a land/sea mask the fused
// momentum stencil can stage as bytes)
inline bool binary_mask(const double *a, size_t n)
{
    for (size_t q = 0; q < n; ++q)
        if (!(a[q] == 0.0 || a[q] == 1.0)) return false;
    return true;
}

// true when the 2D field a (ni x nj, i fastest) is constant along every row
// j (f-plane and beta-plane Coriolis: the fused momentum stencil stages one
// value per row)
inline bool rows_uniform(const double *a, size_t ni, size_t nj)
{
    for (size_t j = 0; j < nj; ++j)
        for (size_t i = 1; i < ni; ++i)
            if (!(a[j * ni + i] == a[j * ni])) return false;
    return true;
}

}  // namespace mom6h
