// mom6h_model.h -- the model object (one run of the split-RK2 driver on one
// GPU), shared by the model and the ensemble sources.
#pragma once

#include <string>
#include <vector>

#include "mom6h_fused.h"

namespace mom6h {

struct Field {
    std::string name;
    void **p;
    size_t n;
    int bytes;    // 8: FP64, 4: FP32 (the layer fields of the mixed-precision form)
    bool owned;   // freed with the field list (BT_cont 2D arrays are freed with their struct)
};

enum Stage {
    ST_TRANSPORTS, ST_HOR_VISC, ST_CORIOLIS, ST_VEL_UPDATE, ST_VERT_VISC, ST_CONTINUITY,
    ST_HALO, ST_BTSTEP, ST_MOMENTUM, ST_COUNT
};
extern const char *kStageNames[ST_COUNT];

}  // namespace mom6h

struct mom6h_model {
    using Dims = mom6h::Dims;
    using Halo = mom6h::Halo;
    using UMetrics = mom6h::UMetrics;
    using Field = mom6h::Field;
    mom6h_config c;
    Dims d;
    Halo halo;
    bool prepared = false;
    int use_graph = 0, profile = 0;
    int momentum_tile = 8;   // rows of the fused momentum tile (4, 8 or 16)
    int momentum_splits = 1; // blocks per tile over the layers (1..nk)
    int serialize = 0;       // barotropic chain on the 3D stream (per-kernel timing without overlap)
    int fused = 1;           // fused momentum stencil (module-by-module when 0)
    bool uniform = false;    // every metric array constant (fused kernels allowed)
    bool fused_rows = false; // masks 0/1 and Coriolis constant along rows (the fused stencil's staging)
    UMetrics ug{};
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t e_start = nullptr, e_bt = nullptr;
    // state and driver arrays
    double *u, *v, *h, *h0, *htmp, *uh, *vh, *CAu, *CAv, *up, *vp, *diffu, *diffv;
    double *eta, *ubt, *vbt, *ubt_av, *vbt_av, *eta_av, *u_cor, *du_cor;
    double *hx = nullptr, *v_cor = nullptr, *dv_cor = nullptr;   // meridional pass
    mom6h_bt_cont bt_v{};                                          // its v-side BT_cont
    double *taux, *tauy, *ray_u, *ray_v, *CoriolisBu, *mask2dT, *mask2dCu, *mask2dCv, *mask2dBu;
    mom6h_bt_cont bt{};
    // mixed-precision form: FP32 layer fields (the FP64 pointers above that
    // name them stay NULL); 2D fields, BT_cont 2D and du_cor stay FP64
    bool mixed = false;
    struct {
        float *u, *v, *h, *h0, *htmp, *uh, *vh, *up, *vp, *u_cor, *bt_h_u;
        float *hx, *v_cor, *bt_h_v;   // meridional pass
        float *del2u, *del2v;         // biharmonic scratch of the fused stencil
    } f{};
    double *del2u = nullptr, *del2v = nullptr;   // biharmonic scratch of the fused stencil (FP64)
    std::vector<Field> fields;
    // host copies of the uniform metrics (init_ocean_grid)
    std::vector<double> IareaT, areaT, dxT, dyT, IdxT, IdyT, dxCu, dyCu, dy_Cu, IdxCu, IdyCu;
    std::vector<double> dxCv, dyCv, IdxCv, IdyCv, IareaBu, IareaCu, IareaCv, dxBu, dyBu, IdxBu, IdyBu;
    // module control structures
    mom6h_continuity_cs *cont = nullptr;
    mom6h_coriolis_cs *cor = nullptr;
    mom6h_barotropic_cs *bt_cs = nullptr;
    mom6h_vert_visc_cs *vv = nullptr;
    mom6h_hor_visc_cs *hv = nullptr;
    // one step captured as two graphs, one per stream (replayed with one fork and one join per call)
    hipGraph_t graph = nullptr;          // one 3D step (stream s0)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph_bt = nullptr;       // one barotropic step (stream s1)
    hipGraphExec_t gexec_bt = nullptr;
    int graph_fused = -1;    // fused_ok() at capture time
    // the coupled form: continuity's barotropic-transport targets (2D)
    double *uhbt = nullptr, *vhbt = nullptr;
    // Rayleigh drag at every level (config ray3d)
    double *ray_u3d = nullptr, *ray_v3d = nullptr;
    // diagnostics (option "diag"): the predictor's vertical sums of diffu*h and
    // diffv*h in one shared 2D work array (the reference's work_2d), KE and
    // mass fields, the reduction workspace and 16 statistics on the device
    int diag = 0;
    double *diag_work2d = nullptr, *diag_ke = nullptr, *diag_mass = nullptr, *diag_red = nullptr;
    double *diag_out = nullptr;
    // profiling
    std::vector<hipEvent_t> ev;
    double stage_ms[mom6h::ST_COUNT] = {0};
};


namespace mom6h {
// the fused stencils apply (uniform metrics, Sadourny Coriolis, Laplacian [+ biharmonic])
bool fused_ok(const mom6h_model *m);
}  // namespace mom6h
