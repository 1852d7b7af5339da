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
};

struct MomArgs {
    Dims d;
    UMetrics g;
    double Kh, h_neglect, fdt;
    int no_slip;
    const double *u, *v, *h;          // stencil inputs
    const double *bu, *bv;            // bases of the update (may alias ou/ov)
    double *ou, *ov;
    const double *mask2dT, *mask2dBu, *CoriolisBu;
};

int launch_momentum(const MomArgs &a, hipStream_t s);

// true and the common value when every element of a[0..n) is identical
inline bool uniform_value(const double *a, size_t n, double *v)
{
    for (size_t q = 1; q < n; ++q)
        if (!(a[q] == a[0])) return false;
    *v = a[0];
    return true;
}

}  // namespace mom6h
